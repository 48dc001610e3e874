// The reference's openQmin driver body compiled unchanged against the drop-in
// C++ classes (open-qmin_amd/host/openqmin_hip.h) and libqmin.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `driver`) copies the body
// of main() in the reference's openQmin.cpp, from the phase constants
// ("scalar a = -1;") to the class-size report, out of the reference checkout
// into oracle/_ref/driver_body.inc, unmodified, and compiles it here into
// oracle/_ref/openQmin-body.  The body #includes the reference's own
// setInitialConditions.h and addObjectsToOpenQmin.h from the checkout.
//
// What precedes the body in the reference (MPI start-up and the TCLAP command
// line) needs MPI and is replaced by this harness: it reads the flags the body
// uses, with the reference's names and defaults, into variables of the
// reference's names, on one rank (rank 0 of a 1x1x1 grid).
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>

#include "openqmin_hip.h"

using namespace std;

namespace {

// the .getValue() of a parsed command-line flag
template <typename T>
struct Flag {
    T value;
    T getValue() const { return value; }
};

struct Args {
    int argc;
    char** argv;
    const char* find(const char* shortName, const char* longName) const {
        for (int k = 1; k + 1 < argc; ++k) {
            const char* a = argv[k];
            if ((shortName && a[0] == '-' && a[1] != '-' && std::strcmp(a + 1, shortName) == 0) ||
                (longName && a[0] == '-' && a[1] == '-' && std::strcmp(a + 2, longName) == 0))
                return argv[k + 1];
        }
        return nullptr;
    }
    bool has(const char* shortName, const char* longName) const {
        for (int k = 1; k < argc; ++k) {
            const char* a = argv[k];
            if ((a[0] == '-' && a[1] != '-' && std::strcmp(a + 1, shortName) == 0) ||
                (a[0] == '-' && a[1] == '-' && std::strcmp(a + 2, longName) == 0))
                return true;
        }
        return false;
    }
    int i(const char* s, const char* l, int d) const {
        const char* v = find(s, l);
        return v ? std::atoi(v) : d;
    }
    double d(const char* s, const char* l, double d0) const {
        const char* v = find(s, l);
        return v ? std::atof(v) : d0;
    }
    std::string str(const char* s, const char* l, const char* d0) const {
        const char* v = find(s, l);
        return v ? std::string(v) : std::string(d0);
    }
};

}  // namespace

int main(int argc, char* argv[]) {
    const Args args{argc, argv};
    int myRank = 0;
    const int3 rankTopology{1, 1, 1};
    string initialConfigurationFile = args.str(nullptr, "initialConfigurationFile", "NONE");
    string fieldFile = args.str(nullptr, "spatiallyVaryingFieldFile", "NONE");
    string boundaryFile = args.str(nullptr, "boundaryFile", "NONE");
    string saveFile = args.str(nullptr, "saveFile", "NONE");
    int saveStride = args.i(nullptr, "stride", 1);
    int linearSave = args.i(nullptr, "linearSpacedSaving", -1);
    scalar logSave = args.d(nullptr, "logSpacedSaving", -1);
    int randomSeed = args.i(nullptr, "randomSeed", -1);
    bool reproducible = true;
    bool verbose = args.has("v", "verbose");
    int initializationSwitch = args.i("z", "initializationSwitch", 0);
    scalar phaseA = args.d("a", "phaseConstantA", -0.172);
    scalar phaseB = args.d("b", "phaseConstantB", -2.12);
    scalar phaseC = args.d("c", "phaseConstantC", 1.73);
    int boxL = args.i("l", "boxL", 50);
    int boxLx = args.i(nullptr, "Lx", 50);
    int boxLy = args.i(nullptr, "Ly", 50);
    int boxLz = args.i(nullptr, "Lz", 50);
    if (boxL != 50) {
        boxLx = boxL;
        boxLy = boxL;
        boxLz = boxL;
    }
    scalar L1 = args.d(nullptr, "L1", 4.64);
    scalar L2 = args.d(nullptr, "L2", 0.);
    scalar L3 = args.d(nullptr, "L3", 0.);
    scalar L4 = args.d(nullptr, "L4", 0.);
    scalar L6 = args.d(nullptr, "L6", 0.);
    scalar dt = args.d("e", "deltaT", 0.0005);
    scalar forceCutoff = args.d("f", "fTarget", 0.000000000001);
    int maximumIterations = args.i("i", "iterations", 100);
    bool GPU = args.i("g", "GPU", -1) >= 0;
    const Flag<scalar> setHFieldXSwitchArg{args.d(nullptr, "hFieldX", 0)};
    const Flag<scalar> setHFieldYSwitchArg{args.d(nullptr, "hFieldY", 0)};
    const Flag<scalar> setHFieldZSwitchArg{args.d(nullptr, "hFieldZ", 0)};
    const Flag<scalar> setHFieldMu0SwitchArg{args.d(nullptr, "hFieldMu0", 1)};
    const Flag<scalar> setHFieldChiSwitchArg{args.d(nullptr, "hFieldChi", 1)};
    const Flag<scalar> setHFieldDeltaChiSwitchArg{args.d(nullptr, "hFieldDeltaChi", 0.5)};
    const Flag<scalar> setEFieldXSwitchArg{args.d(nullptr, "eFieldX", 0)};
    const Flag<scalar> setEFieldYSwitchArg{args.d(nullptr, "eFieldY", 0)};
    const Flag<scalar> setEFieldZSwitchArg{args.d(nullptr, "eFieldZ", 0)};
    const Flag<scalar> setEFieldEps0SwitchArg{args.d(nullptr, "eFieldEpsilon0", 1)};
    const Flag<scalar> setEFieldEpsSwitchArg{args.d(nullptr, "eFieldEpsilon", 1)};
    const Flag<scalar> setEFieldDeltaEpsSwitchArg{args.d(nullptr, "eFieldDeltaEpsilon", 0.5)};
    try {
#include "driver_body.inc"
    } catch (const qminError& e) {
        std::fprintf(stderr, "qmin error: %s\n", e.what());
        return 1;
    }
    return 0;
}
