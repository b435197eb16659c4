// Command line of the reference (ZQ_SmokeSimulation3D/main.cpp:4-21):
//   ZQ_SmokeSimulation3D <config> <output folder>
#include <cstdio>

#include "ZQ_SmokeSimulation3D.h"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::printf(".exe configfile outfold\n");
        return 1;
    }
    ZQ_SmokeSimulation3D sim;
    if (!sim.Init(argv[1])) {
        std::printf("failed to init %s\n", argv[1]);
        return 1;
    }
    if (!sim.Run(argv[2])) {
        std::printf("failed to run\n");
        return 1;
    }
    return 0;
}
