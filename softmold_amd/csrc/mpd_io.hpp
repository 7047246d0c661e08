// Host-side `.mpd` system description, reader, writer and the two benchmark generators.
//
// The `.mpd` format is the reference's keyword script: whitespace-separated tokens, a
// command name followed by its values, in the command order of include/system.h:292-304.
// Reader semantics follow Blob::input (system.h:589-1313) and Script::read
// (scriptFormat.h:79-94); the writer reproduces Blob::output + Script::write
// (system.h:1315-1749, scriptFormat.h:63-77) byte for byte.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace smio {

// Command ids in the reference's order (system.h:292-304).
enum Cmd {
    GAMMA, INITIALTEMP, FINALTEMP, SEED, NTYPES, NMOLECULES, NPARTICLES, PERIODIC, CUTOFF,
    SIZE, INITIALTIME, FINALTIME, DELTAT, STOREINTERVAL, MEASUREINTERVAL, TWOBODYFCONST,
    TWOBODYUCONST, POSITIONS, VELOCITIES, MOLECULE, BANANA, DELTALXY, REMOVESOLVENT,
    TEMPSTEPINTERVAL, SOLVENTGAMMA, GAMMATYPE, TENSION, NCMD
};
extern const char *const kCmdName[NCMD];

// Molecule type ids (include/algorithms/molecules.h:6-31).
enum MolType {
    CHAINOLD = 0, BENDOLD = 1, ALLPAIRS = 2, TORSION = 3, BONDOLD = 4, DIHEDRAL = 5, BOND = 6,
    BEND = 7, CHAIN = 8, BEAD = 9, SOLID = 10, BOUNDARY = 11, RIGIDBEND = 12, PULLBEAD = 13,
    OFFSET_BOUNDARY = 14, FLOATING_BASE = 15, ZTORQUE = 16, ZPOWERPOTENTIAL = 17,
    NANOCORE = 18, BALL = 19
};

struct Molecule {
    int type = 0;
    std::vector<double> constants;
    std::vector<std::array<int, 4>> bonds;
};

struct System {
    bool present[NCMD] = {};
    double gamma = 0, initialTemp = 0, finalTemp = 0;
    int seed = 0, nTypes = 0, nMolecules = 0, nParticles = 0;
    bool periodic = true;
    double cutoff = 0;
    double size[3] = {0, 0, 0};
    double initialTime = 0, finalTime = 0, deltaT = 0, storeInterval = 0, measureInterval = 0;
    std::vector<double> fconst, uconst;
    std::vector<int> type;
    std::vector<double> pos, vel;        // n*3
    std::vector<Molecule> mol;
    double deltaLXY = 0, removeSolvent = 0, tempStepInterval = 0;
    double solventGamma[2] = {0, 0};
    std::vector<double> gammaType;
    double tension = 0;
};

// Constants and particle indices per record for a molecule type (system.h:1036-1120).
void molecule_layout(int type, int nTypes, int nBondsAlloc, int *nConst, int *nPart);

// Throws std::runtime_error with a message on failure.
System read_file(const std::string &path);
System read_text(const std::string &text);
std::string write_text(const System &s);
void write_file(const System &s, const std::string &path);
// Blob::errorChecking (system.h:438-545).
void check(const System &s);

// Generators restated from generate/liposome.cpp:1-97 + models/lipidModels.h:1313-1469 and
// generate/bilayer.cpp:1-125 (no anchors) + models/lipidModels.h:1740-1924.
System gen_liposome(int seed, int nLipids, double arealDensity);
System gen_bilayer(int seed, int nLipids, double arealDensity);

}  // namespace smio
