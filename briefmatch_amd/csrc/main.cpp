// briefmatch: the command-line tool (the reference's main.cpp:44-89) on the
// MI355X engine.  Exit status 1 on any error.
#include <cstdio>
#include <exception>

#include "../../include/of_pipeline.h"

int main(int argc, char *argv[])
{
    OFPipeline of;
    try {
        if (!of.setParams(argc, argv)) return 1;
        of.setup();
        std::printf("\n\nRunning optical flow loop...\n");
        while (of.processFrame()) {
        }
        std::printf("done!\n");
    } catch (BriefMatchException &e) {
        std::fprintf(stderr, "\nbriefmatch error: %s\n", e.what());
        return 1;
    } catch (std::exception &e) {
        std::fprintf(stderr, "\nError: %s\n", e.what());
        return 1;
    }
    return 0;
}
