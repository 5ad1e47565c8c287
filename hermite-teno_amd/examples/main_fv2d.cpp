// main.cpp-style driver: the reference's driver (main.cpp:1-176) selects one
// case by commenting code; this one takes the case name as argv[1]
// (vortex, entropy, riemann on FV::System2D; hvortex, hentropy, hriemann on
// FVHermite::System2D; sod, entropy1d on FV::System1D).  It
// compiles against include/ and links libhtx.so; the class names and calls
// are the reference's.
#include <cstdio>
#include <cstring>
#include <string>

#include "FVHermite_System2D.h"
#include "FV_System1D.h"
#include "FV_System2D.h"

int main(int argc, char** argv) {
    std::string which = argc > 1 ? argv[1] : "vortex";
    int n = argc > 2 ? std::atoi(argv[2]) : 64;
    try {
        if (which == "vortex") {
            // reference main.cpp:170-173 form, on FV::System2D: vortex at the origin
            FV::System2D systemObj(-5.0, 5.0, n, -5.0, 5.0, n, 0.4, 0.05, 0.0, 1.4);
            systemObj.isentropicVortex();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "entropy") {
            // reference main.cpp:139-142 (commented case)
            FV::System2D systemObj(-1.0, 1.0, n, -1.0, 1.0, n, 0.6, 0.05, 0.0, 1.4);
            systemObj.entropyWave();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "riemann") {
            // reference main.cpp:128-131 (commented case), HLLC variant
            htx::Options o;
            o.flux = HTX_FLUX_HLLC;
            FV::System2D systemObj(0.0, 1.0, n, 0.0, 1.0, n, 0.4, 0.02, 0.0, 1.4, o);
            systemObj.Riemann2D();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "hvortex") {
            // reference main.cpp:142-145 form (commented case), vortex at the origin
            FVHermite::System2D systemObj(-5.0, 5.0, n, -5.0, 5.0, n, 0.4, 0.05, 0.0, 1.4);
            systemObj.isentropicVortex();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "hentropy") {
            // reference main.cpp:102-105 (commented case)
            FVHermite::System2D systemObj(-1.0, 1.0, n, -1.0, 1.0, n, 0.4, 0.05, 0.0, 1.4);
            systemObj.entropyWave();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "hriemann") {
            // reference main.cpp:122-125 (commented case), configuration 5
            FVHermite::System2D systemObj(0.0, 1.0, n, 0.0, 1.0, n, 0.4, 0.02, 0.0, 1.4);
            systemObj.Riemann2D();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "sod") {
            // reference main.cpp:55-58 form (commented case) as the 1D
            // configuration: Sod, TENO5 + HLLC, CFL 0.5 to t = 0.2
            htx::Options1D o;
            o.scheme = HTX_SCHEME_TENO5;
            o.flux = HTX_FLUX_HLLC;
            FV::System1D systemObj(0.0, 1.0, n, 0.5, 0.2, 0.0, 1.4, o);
            systemObj.sodShock();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else if (which == "entropy1d") {
            FV::System1D systemObj(-1.0, 1.0, n, 0.4, 0.5, 0.0, 1.4);
            systemObj.entropyWave();
            systemObj.timeStepping();
            systemObj.printResult();
            std::printf("steps %d t %.17g\n", systemObj.steps(), systemObj.time());
        } else {
            std::fprintf(stderr, "unknown case %s\n", which.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
