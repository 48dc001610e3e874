// openQmin-hip: the reference's openQmin driver (openQmin.cpp) for one MI355X,
// written against the drop-in classes of openqmin_hip.h.
//
// Flags keep the reference's names and defaults (openQmin.cpp:59-115):
//   -l/--boxL, --Lx/--Ly/--Lz, -i/--iterations, -f/--fTarget, -e/--deltaT,
//   -a/-b/-c phase constants, --L1 --L2 --L3 --L4 --L6, --randomSeed,
//   --boundaryFile, --saveFile, --linearSpacedSaving, --logSpacedSaving, --stride,
//   --initialConfigurationFile, --spatiallyVaryingFieldFile,
//   --hField{X,Y,Z,Mu0,Chi,DeltaChi},
//   --eField{X,Y,Z,Epsilon0,Epsilon,DeltaEpsilon}, -z/--initializationSwitch (-1, 0, 1, 2, 3 as
//   setInitialConditions.h:21-99, with its example director function for 3), -v.
// Additions of this build (no reference flag):
//   --minimizer fire|adam|nesterov|lolbfgs   (default fire, as openQmin)
//   --precision fp64|fp32|mixed              (mixed = FP32 to the tolerance, then FP64 polish)
//   --wall xyz,plane,type,W                  (type 0 homeotropic, 1 degenerate planar; S0 from the phase)
//   --sphere cx,cy,cz,r,type,W
//   --cpuSemantics                           (the reference CPU path's double metric correction)
// Output: one summary line "iterations <n> forceMax <f> energy <E/N_active> seconds <t>".
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>

#include "openqmin_hip.h"

namespace {

struct Args {
    std::map<std::string, std::string> v;
    std::vector<std::string> walls, spheres;
    bool has(const std::string& k) const { return v.count(k) != 0; }
    double d(const std::string& k, double def) const { return has(k) ? std::atof(v.at(k).c_str()) : def; }
    int i(const std::string& k, int def) const { return has(k) ? std::atoi(v.at(k).c_str()) : def; }
    std::string s(const std::string& k, const std::string& def) const { return has(k) ? v.at(k) : def; }
};

const std::map<std::string, std::string> kShort = {
    {"-l", "boxL"}, {"-i", "iterations"}, {"-f", "fTarget"}, {"-e", "deltaT"}, {"-a", "phaseConstantA"},
    {"-b", "phaseConstantB"}, {"-c", "phaseConstantC"}, {"-z", "initializationSwitch"}, {"-g", "GPU"}};
const char* kSwitches[] = {"-v", "--verbose", "-r", "--reproducible", "--cpuSemantics"};
const char* kValued[] = {"boxL", "Lx", "Ly", "Lz", "iterations", "fTarget", "deltaT", "phaseConstantA",
                         "phaseConstantB", "phaseConstantC", "initializationSwitch", "GPU", "randomSeed", "L1", "L2",
                         "L3", "L4", "L6", "initialConfigurationFile", "spatiallyVaryingFieldFile", "boundaryFile",
                         "saveFile", "linearSpacedSaving", "logSpacedSaving", "stride", "hFieldX", "hFieldY",
                         "hFieldZ", "hFieldMu0", "hFieldChi", "hFieldDeltaChi", "eFieldX", "eFieldY", "eFieldZ",
                         "eFieldEpsilon0", "eFieldEpsilon", "eFieldDeltaEpsilon", "minimizer", "precision", "wall",
                         "sphere", "lbfgsMemory"};

std::vector<double> numbers(const std::string& csv, size_t n, const char* what) {
    std::vector<double> out;
    std::stringstream ss(csv);
    std::string item;
    while (std::getline(ss, item, ',')) out.push_back(std::atof(item.c_str()));
    if (out.size() != n) throw std::runtime_error(std::string("--") + what + " takes " + std::to_string(n) + " values");
    return out;
}

Args parse(int argc, char** argv) {
    Args a;
    for (int k = 1; k < argc; ++k) {
        std::string key = argv[k];
        bool isSwitch = false;
        for (const char* s : kSwitches) isSwitch |= key == s;
        if (isSwitch) {
            a.v[key == "-v" ? "verbose" : key.substr(key.find_first_not_of('-'))] = "1";
            continue;
        }
        if (k + 1 >= argc) throw std::runtime_error("missing value for " + key);
        std::string val = argv[++k];
        if (kShort.count(key))
            key = kShort.at(key);
        else if (key.rfind("--", 0) == 0)
            key = key.substr(2);
        else
            throw std::runtime_error("unknown flag " + key);
        bool known = false;
        for (const char* name : kValued) known |= key == name;
        if (!known) throw std::runtime_error("unknown flag --" + key);
        if (key == "minimizer" && val != "fire" && val != "adam" && val != "nesterov" && val != "lolbfgs")
            throw std::runtime_error("unknown --minimizer " + val);
        if (key == "wall") {
            numbers(val, 4, "wall");
            a.walls.push_back(val);
        } else if (key == "sphere") {
            numbers(val, 6, "sphere");
            a.spheres.push_back(val);
        }
        else
            a.v[key] = val;
    }
    if (a.has("initializationSwitch") && (a.i("initializationSwitch", 0) < -1 || a.i("initializationSwitch", 0) > 3))
        throw std::runtime_error("-z is -1 (file), 0 (random), 1 (random uniform), 2 (fixed uniform) or 3 (function)");
    if (a.has("initializationSwitch") && a.i("initializationSwitch", 0) == -1 &&
        a.s("initialConfigurationFile", "NONE") == "NONE")
        throw std::runtime_error("-z -1 needs --initialConfigurationFile");
    if (a.has("precision") && a.v.at("precision") != "fp64" && a.v.at("precision") != "fp32" &&
        a.v.at("precision") != "mixed")
        throw std::runtime_error("--precision is fp64, fp32 or mixed");
    if (a.s("precision", "fp64") == "mixed" && (a.i("linearSpacedSaving", -1) > 0 || a.d("logSpacedSaving", -1) > 0))
        throw std::runtime_error("--precision mixed does not combine with save schedules");
    return a;
}


}  // namespace

int main(int argc, char** argv) {
    try {
        const Args a = parse(argc, argv);
        const bool verbose = a.has("verbose");
        int boxL = a.i("boxL", 50);
        const int lx = a.has("Lx") ? a.i("Lx", 50) : boxL;
        const int ly = a.has("Ly") ? a.i("Ly", 50) : boxL;
        const int lz = a.has("Lz") ? a.i("Lz", 50) : boxL;
        const double phaseA = a.d("phaseConstantA", -0.172), phaseB = a.d("phaseConstantB", -2.12),
                     phaseC = a.d("phaseConstantC", 1.73);
        const double L1 = a.d("L1", 4.64), L2 = a.d("L2", 0), L3 = a.d("L3", 0), L4 = a.d("L4", 0),
                     L6 = a.d("L6", 0);
        const double dt = a.d("deltaT", 0.0005), forceCutoff = a.d("fTarget", 1e-12);
        const int maximumIterations = a.i("iterations", 100);
        const int randomSeed = a.i("randomSeed", -1);
        const std::string minimizer = a.s("minimizer", "fire");
        const std::string precision = a.s("precision", "fp64");
        if (precision != "fp64" && precision != "fp32" && precision != "mixed")
            throw std::runtime_error("--precision is fp64, fp32 or mixed");

        // openQmin.cpp:173-180, 187-199
        const scalar A = -1, B = -phaseB / phaseA, C = -phaseC / phaseA;
        noiseSource noise(true);
        noise.setReproducibleSeed(randomSeed == -1 ? 13371 : randomSeed);
        const bool oneConstant = !(L2 != 0. || L3 != 0. || L4 != 0. || L6 != 0.);
        auto Configuration = std::make_shared<multirankQTensorLatticeModel>(
            lx, ly, lz, false, false, false, false, false, precision == "fp64" ? QM_FP64 : QM_FP32);
        auto sim = std::make_shared<multirankSimulation>(0, 1, 1, 1, false, false);
        auto landauLCForce = std::make_shared<landauDeGennesLC>();
        sim->setConfiguration(Configuration);
        landauLCForce->setPhaseConstants(A, B, C);
        if (oneConstant) {
            landauLCForce->setElasticConstants(L1);
            landauLCForce->setNumberOfConstants(distortionEnergyType::oneConstant);
        } else {
            landauLCForce->setElasticConstants(L1, L2, L3, L4, L6);
            landauLCForce->setNumberOfConstants(distortionEnergyType::multiConstant);
        }
        const scalar3 H = make_scalar3(a.d("hFieldX", 0), a.d("hFieldY", 0), a.d("hFieldZ", 0));
        const scalar3 E = make_scalar3(a.d("eFieldX", 0), a.d("eFieldY", 0), a.d("eFieldZ", 0));
        if (H.x != 0 || H.y != 0 || H.z != 0)
            landauLCForce->setHField(H, a.d("hFieldChi", 1), a.d("hFieldMu0", 1), a.d("hFieldDeltaChi", 0.5));
        if (E.x != 0 || E.y != 0 || E.z != 0)
            landauLCForce->setEField(E, a.d("eFieldEpsilon", 1), a.d("eFieldEpsilon0", 1),
                                     a.d("eFieldDeltaEpsilon", 0.5));
        landauLCForce->setModel(Configuration);
        sim->addForce(landauLCForce);
        // openQmin.cpp:245-248
        if (a.s("spatiallyVaryingFieldFile", "NONE") != "NONE")
            landauLCForce->setSpatiallyVaryingField(a.s("spatiallyVaryingFieldFile", ""), a.d("hFieldChi", 1),
                                                    a.d("hFieldMu0", 1), a.d("hFieldDeltaChi", 0.5));
        sim->setCPUOperation(a.has("cpuSemantics"));

        std::shared_ptr<updater> upd;
        if (minimizer == "fire") {
            auto m = std::make_shared<energyMinimizerFIRE>(Configuration);
            // openQmin.cpp:254-258
            m->setFIREParameters(dt, .99, 100 * dt, 1.1, 0.95, 0.9, 4, forceCutoff, 0.0);
            upd = m;
        } else if (minimizer == "adam") {
            auto m = std::make_shared<energyMinimizerAdam>(Configuration);
            m->setAdamParameters(0.9, 0.99, 1e-8, dt, forceCutoff);
            upd = m;
        } else if (minimizer == "nesterov") {
            auto m = std::make_shared<energyMinimizerNesterovAG>(Configuration);
            m->setNesterovAGParameters(dt, 0.01, forceCutoff);
            upd = m;
        } else if (minimizer == "lolbfgs") {
            auto m = std::make_shared<energyMinimizerLoLBFGS>(Configuration);
            m->setLoLBFGSParameters(a.i("lbfgsMemory", 5), dt, 1.0, forceCutoff, 10.0);
            upd = m;
        } else {
            throw std::runtime_error("unknown --minimizer " + minimizer);
        }
        upd->setMaximumIterations(maximumIterations);
        sim->addUpdater(upd, Configuration);

        const scalar S0 = (-B + std::sqrt(B * B - 24 * A * C)) / (6 * C);
        // setInitialConditions.h:21-99
        const std::string init = a.s("initialConfigurationFile", "NONE");
        const int initializationSwitch = init != "NONE" ? -1 : a.i("initializationSwitch", 0);
        if (initializationSwitch == -1) {
            sim->loadState(init);
        } else if (initializationSwitch == 0) {
            Configuration->setNematicQTensorRandomly(noise, S0);
        } else if (initializationSwitch == 1) {
            // one random director for every site, from its own reproducible stream
            noiseSource multirankNoise(true);
            multirankNoise.setReproducibleSeed(randomSeed == -1 ? 13371 : randomSeed);
            Configuration->setRandomDirectors(multirankNoise, S0, true);
        } else if (initializationSwitch == 2) {
            scalar3 targetDirector = make_scalar3(1. / 2, 0, 1 / 2.);
            normalizeDirector(targetDirector);
            Configuration->setUniformDirectors(targetDirector, S0);
        } else {
            // the example function of setInitialConditions.h
            Configuration->setDirectorFromFunction([](scalar x, scalar y, scalar z) {
                scalar3 director = make_scalar3(x, x + y + std::cos(z), 0);
                const scalar s0 = 0.53 + 0.0000000001 * (x + y + z);
                normalizeDirector(director);
                return scalar4{director.x, director.y, director.z, s0};
            });
        }

        if (a.s("boundaryFile", "NONE") != "NONE") sim->createBoundaryFromFile(a.s("boundaryFile", ""), verbose);
        for (const std::string& w : a.walls) {
            const auto p = numbers(w, 4, "wall");
            boundaryObject b(p[2] == 0 ? boundaryType::homeotropic : boundaryType::degeneratePlanar, p[3], S0);
            sim->createWall(int(p[0]), int(p[1]), b);
        }
        for (const std::string& s : a.spheres) {
            const auto p = numbers(s, 6, "sphere");
            boundaryObject b(p[4] == 0 ? boundaryType::homeotropic : boundaryType::degeneratePlanar, p[5], S0);
            sim->createSphericalColloid(make_scalar3(p[0], p[1], p[2]), p[3], b);
        }
        sim->finalizeObjects();

        const auto t0 = std::chrono::steady_clock::now();
        // openQmin.cpp:291-333: minimize, or minimize in stages saving the state
        // before each (file names carry the iteration, "_t<iteration>")
        const int linearSave = a.i("linearSpacedSaving", -1);
        const scalar logSave = a.d("logSpacedSaving", -1);
        const int saveStride = a.i("stride", 1);
        const std::string saveFile = a.s("saveFile", "NONE");
        if (linearSave < 0 && logSave < 0) {
            sim->performTimestep();
        } else if (logSave < 0) {
            int currentIteration = 0;
            while (currentIteration + linearSave < maximumIterations && upd->getMaxForce() > forceCutoff) {
                const std::string newSaveFile = saveFile + "_t" + std::to_string(currentIteration);
                if (saveFile != "NONE") sim->saveState(newSaveFile, saveStride);
                currentIteration += linearSave;
                upd->setMaximumIterations(currentIteration);
                sim->performTimestep();
                if (verbose) std::printf("saving to %s (plus _xX_yY_zZ.txt)\n", newSaveFile.c_str());
            }
        } else {
            int currentIteration = 0;
            logSpacedIntegers lsi(0, logSave);
            while (currentIteration < maximumIterations && upd->getMaxForce() > forceCutoff) {
                const std::string newSaveFile = saveFile + "_t" + std::to_string(currentIteration);
                if (saveFile != "NONE") sim->saveState(newSaveFile, saveStride);
                lsi.update();
                currentIteration = lsi.nextSave;
                upd->setMaximumIterations(currentIteration);
                sim->performTimestep();
                if (verbose) std::printf("saving to %s (plus _xX_yY_zZ.txt)\n", newSaveFile.c_str());
            }
        }
        if (precision == "mixed") {
            // FP64 polish: the state converted in place, the minimizer continues
            // from where the FP32 phase stopped
            Configuration->setPrecision(QM_FP64);
            upd->setMaximumIterations(upd->iterations + maximumIterations);
            sim->performTimestep();
        }
        const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const scalar e = sim->computePotentialEnergy(verbose);
        std::printf("iterations %d forceMax %.17g energy %.17g seconds %.6f\n", upd->iterations, upd->getMaxForce(),
                    e, seconds);
        if (saveFile != "NONE") sim->saveState(saveFile, saveStride);
        return 0;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "openQmin-hip: %s\n", ex.what());
        return 1;
    }
}
