// `.mpd` reader/writer and the liposome/bilayer generators (host C++).
#include "mpd_io.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>

namespace smio {

const char *const kCmdName[NCMD] = {
    "gamma", "initialTemp", "finalTemp", "seed", "nTypes", "nMolecules", "nParticles",
    "periodic", "cutoff", "size", "initialTime", "finalTime", "deltaT", "storeInterval",
    "measureInterval", "twoBodyFconst", "twoBodyUconst", "positions", "velocities",
    "molecule", "banana", "deltaLXY", "removeSolvent", "tempStepInterval", "solventGamma",
    "gammaType", "tension"};

static const int kBeadConst = 22;  // nBEADCONST (include/MD.h:52)

void molecule_layout(int type, int nTypes, int nBondsAlloc, int *nConst, int *nPart)
{
    int c = 0, p = 0;
    switch (type) {
    case BOND: c = 2; p = 2; break;
    case BEND: c = 2; p = 3; break;
    case CHAIN: c = 4; p = 3; break;
    case BEAD: c = kBeadConst * nTypes * nTypes; p = 1; break;
    case SOLID: c = 1; p = 1; break;
    case BOUNDARY: c = 4; p = 1; break;
    case OFFSET_BOUNDARY: c = 4; p = 1; break;
    case RIGIDBEND: c = 5; p = 2; break;
    case PULLBEAD: c = 4; p = 1; break;
    case FLOATING_BASE: c = 6 * nTypes; p = 1; break;
    case ZTORQUE: c = 4; p = 3; break;
    case ZPOWERPOTENTIAL: c = 2; p = 2; break;
    case NANOCORE: c = kBeadConst * nBondsAlloc; p = 1; break;
    case BALL: c = 2; p = 2; break;
    default: c = 0; p = 0; break;  // TORSION, DIHEDRAL, old types: 4 raw ints
    }
    *nConst = c;
    *nPart = p;
}

// ---------------------------------------------------------------------------------------
// reader
// ---------------------------------------------------------------------------------------
namespace {

struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    bool more() const { return i < t.size(); }
    const std::string &next(const char *what)
    {
        if (i >= t.size())
            throw std::runtime_error(std::string("unexpected end of file while reading ") + what);
        return t[i++];
    }
};

[[noreturn]] void fail_num(const char *key, const std::string &tok)
{
    throw std::runtime_error(std::string("Value after ") + key + " is not a numerical type: '" + tok + "'");
}

// stringstream >> double reads a numeric prefix and fails when there is none.
double parse_f(const char *key, const std::string &tok)
{
    const char *s = tok.c_str();
    char *end = nullptr;
    errno = 0;
    double v = std::strtod(s, &end);
    if (end == s || std::isnan(v) || std::isinf(v)) fail_num(key, tok);
    return v;
}

int parse_i(const char *key, const std::string &tok)
{
    const char *s = tok.c_str();
    char *end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s) fail_num(key, tok);
    return (int)v;
}

}  // namespace

System read_text(const std::string &text)
{
    Tokens T;
    {
        // Script::read: tokens until EOF or the literal "end"; a last token that is not
        // followed by whitespace hits EOF and is dropped (scriptFormat.h:86-91).
        std::istringstream in(text);
        std::string w;
        while (in >> w) {
            if (in.eof()) break;
            if (w == "end") break;
            T.t.push_back(w);
        }
    }
    System s;
    auto need = [&](Cmd c, const char *what) {
        if (!s.present[c])
            throw std::runtime_error(std::string(kCmdName[c]) + " was not present before " + what + "!");
    };
    while (T.more()) {
        const std::string &key = T.next("a command");
        int c = -1;
        for (int k = 0; k < NCMD; k++)
            if (key == kCmdName[k]) { c = k; break; }
        if (c < 0) throw std::runtime_error(key + " is not a recognized command!");
        const char *kn = kCmdName[c];
        switch (c) {
        case GAMMA: s.gamma = parse_f(kn, T.next(kn)); break;
        case INITIALTEMP: s.initialTemp = parse_f(kn, T.next(kn)); break;
        case FINALTEMP: s.finalTemp = parse_f(kn, T.next(kn)); break;
        case SEED: s.seed = parse_i(kn, T.next(kn)); break;
        case NTYPES: s.nTypes = parse_i(kn, T.next(kn)); break;
        case NMOLECULES: s.nMolecules = parse_i(kn, T.next(kn)); break;
        case NPARTICLES: s.nParticles = parse_i(kn, T.next(kn)); break;
        case PERIODIC: T.next(kn); s.periodic = true; break;  // any value: all axes periodic (system.h:745-754)
        case CUTOFF: s.cutoff = parse_f(kn, T.next(kn)); break;
        case SIZE:
            for (int d = 0; d < 3; d++) s.size[d] = parse_f(kn, T.next(kn));
            break;
        case INITIALTIME: s.initialTime = parse_f(kn, T.next(kn)); break;
        case FINALTIME: s.finalTime = parse_f(kn, T.next(kn)); break;
        case DELTAT: s.deltaT = parse_f(kn, T.next(kn)); break;
        case STOREINTERVAL: s.storeInterval = parse_f(kn, T.next(kn)); break;
        case MEASUREINTERVAL: s.measureInterval = parse_f(kn, T.next(kn)); break;
        case TWOBODYFCONST:
        case TWOBODYUCONST: {
            need(NTYPES, kn);
            std::vector<double> &v = (c == TWOBODYFCONST) ? s.fconst : s.uconst;
            v.assign(6 * (size_t)s.nTypes * s.nTypes, 0.0);
            for (auto &x : v) x = parse_f(kn, T.next(kn));
            break;
        }
        case POSITIONS: {
            need(NPARTICLES, kn);
            s.type.assign(s.nParticles, 0);
            s.pos.assign(3 * (size_t)s.nParticles, 0.0);
            for (int i = 0; i < s.nParticles; i++) {
                s.type[i] = parse_i(kn, T.next(kn));
                for (int d = 0; d < 3; d++) s.pos[3 * (size_t)i + d] = parse_f(kn, T.next(kn));
            }
            break;
        }
        case VELOCITIES: {
            need(NPARTICLES, kn);
            s.vel.assign(3 * (size_t)s.nParticles, 0.0);
            for (auto &x : s.vel) x = parse_f(kn, T.next(kn));
            break;
        }
        case MOLECULE: {
            need(NMOLECULES, kn);
            s.mol.clear();
            for (int m = 0; m < s.nMolecules; m++) {
                Molecule mo;
                mo.type = parse_i(kn, T.next(kn));
                int nc, np;
                molecule_layout(mo.type, s.nTypes, 0, &nc, &np);
                if (mo.type == BEAD && s.nTypes <= 0)
                    throw std::runtime_error("nTypes undefined for BEAD molecule!");
                if (np == 0) {
                    std::array<int, 4> b{};
                    for (int k = 0; k < 4; k++) b[k] = parse_i(kn, T.next(kn));
                    mo.bonds.push_back(b);
                } else {
                    int nb = parse_i(kn, T.next(kn));
                    if (nb <= 0) throw std::runtime_error("molecule record with no bonds is not supported");
                    molecule_layout(mo.type, s.nTypes, nb, &nc, &np);
                    mo.constants.resize(nc);
                    for (auto &x : mo.constants) x = parse_f(kn, T.next(kn));
                    mo.bonds.resize(nb);
                    for (int b = 0; b < nb; b++) {
                        mo.bonds[b] = {0, 0, 0, 0};
                        for (int k = 0; k < np; k++) mo.bonds[b][k] = parse_i(kn, T.next(kn));
                    }
                }
                s.mol.push_back(mo);
            }
            break;
        }
        case BANANA: break;
        case DELTALXY: s.deltaLXY = parse_f(kn, T.next(kn)); break;
        case REMOVESOLVENT: s.removeSolvent = parse_f(kn, T.next(kn)); break;
        case TEMPSTEPINTERVAL: s.tempStepInterval = parse_f(kn, T.next(kn)); break;
        case SOLVENTGAMMA:
            for (int k = 0; k < 2; k++) s.solventGamma[k] = parse_f(kn, T.next(kn));
            break;
        case GAMMATYPE:
            need(NTYPES, kn);
            s.gammaType.assign(s.nTypes, 0.0);
            for (auto &x : s.gammaType) x = parse_f(kn, T.next(kn));
            break;
        case TENSION: s.tension = parse_f(kn, T.next(kn)); break;
        }
        s.present[c] = true;
    }
    check(s);
    return s;
}

System read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Could not open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return read_text(ss.str());
}

void check(const System &s)
{
    const size_t n = (size_t)s.nParticles;
    if (s.pos.size() != 3 * n || s.vel.size() != 3 * n || s.type.size() != n)
        throw std::runtime_error("Mismatch of parameters: nParticles " + std::to_string(n) +
                                 ", positions " + std::to_string(s.type.size()) +
                                 ", velocities " + std::to_string(s.vel.size() / 3));
    const char ax[3] = {'X', 'Y', 'Z'};
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double p = s.pos[3 * i + d];
            if (p > s.size[d] || p < 0)
                throw std::runtime_error(std::string(1, ax[d]) + " position of particle " +
                                         std::to_string(i) + " is out of bounds.");
        }
    for (size_t m = 0; m < s.mol.size(); m++) {
        const Molecule &mo = s.mol[m];
        for (const auto &b : mo.bonds) {
            if (mo.type == CHAIN) {
                long end = (long)b[0] + (long)b[1] * b[2];
                if (end > (long)n || end < 0)
                    throw std::runtime_error("CHAIN Molecule " + std::to_string(m) + " is out of bounds!");
            } else if (mo.type == BOND || mo.type == BEND) {
                int np = (mo.type == BOND) ? 2 : 3;
                for (int k = 0; k < np; k++)
                    if (b[k] > (int)n || b[k] < 0)
                        throw std::runtime_error("Molecule " + std::to_string(m) + " index out of bounds!");
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// writer: the token stream of Blob::output joined by Script::write, i.e. every token
// followed by one space, absent commands contributing an empty token, and a final '\n'.
// Numbers use setprecision(15) on a default-format stream, i.e. %.15g.
// ---------------------------------------------------------------------------------------
namespace {

struct Out {
    std::string s;
    char buf[64];
    void tok(const std::string &t) { s += t; s += ' '; }
    std::string f(double v)
    {
        std::snprintf(buf, sizeof buf, "%.15g", v);
        return buf;
    }
    std::string i(long v) { return std::to_string(v); }
};

}  // namespace

std::string write_text(const System &s)
{
    Out o;
    o.s.reserve(64 * (size_t)s.nParticles + 4096);
    for (int c = 0; c < NCMD; c++) {
        if (!s.present[c]) {
            o.tok("");
            continue;
        }
        o.tok(std::string(c ? "\n" : "") + kCmdName[c]);
        switch (c) {
        case GAMMA: o.tok(o.f(s.gamma)); break;
        case INITIALTEMP: o.tok(o.f(s.initialTemp)); break;
        case FINALTEMP: o.tok(o.f(s.finalTemp)); break;
        case SEED: o.tok(o.i(s.seed)); break;
        case NTYPES: o.tok(o.i(s.nTypes)); break;
        case NMOLECULES: o.tok(o.i(s.nMolecules)); break;
        case NPARTICLES: o.tok(o.i(s.nParticles)); break;
        case PERIODIC: o.tok(s.periodic ? "1" : "0"); break;
        case CUTOFF: o.tok(o.f(s.cutoff)); break;
        case SIZE: for (int d = 0; d < 3; d++) o.tok(o.f(s.size[d])); break;
        case INITIALTIME: o.tok(o.f(s.initialTime)); break;
        case FINALTIME: o.tok(o.f(s.finalTime)); break;
        case DELTAT: o.tok(o.f(s.deltaT)); break;
        case STOREINTERVAL: o.tok(o.f(s.storeInterval)); break;
        case MEASUREINTERVAL: o.tok(o.f(s.measureInterval)); break;
        case TWOBODYFCONST:
        case TWOBODYUCONST: {
            const std::vector<double> &v = (c == TWOBODYFCONST) ? s.fconst : s.uconst;
            for (size_t it = 1; it <= v.size(); it++) {
                std::string t = (it == 1) ? "\n " : "";
                t += o.f(v[it - 1]);
                if (it % 6 == 0) t += "\n";
                o.tok(t);
            }
            break;
        }
        case POSITIONS:
            for (size_t i = 0; i < (size_t)s.nParticles; i++) {
                o.tok((i == 0 ? "\n " : "") + o.i(s.type[i]));
                o.tok(o.f(s.pos[3 * i]));
                o.tok(o.f(s.pos[3 * i + 1]));
                o.tok(o.f(s.pos[3 * i + 2]) + "\n");
            }
            break;
        case VELOCITIES:
            for (size_t i = 0; i < (size_t)s.nParticles; i++) {
                o.tok((i == 0 ? "\n " : "") + o.f(s.vel[3 * i]));
                o.tok(o.f(s.vel[3 * i + 1]));
                o.tok(o.f(s.vel[3 * i + 2]) + "\n");
            }
            break;
        case MOLECULE:
            for (size_t m = 0; m < s.mol.size(); m++) {
                const Molecule &mo = s.mol[m];
                o.tok((m == 0 ? "\n" : "") + o.i(mo.type) + "\t");
                int nc, np;
                molecule_layout(mo.type, s.nTypes, (int)mo.bonds.size(), &nc, &np);
                if (np == 0) {
                    for (int k = 0; k < 4; k++) o.tok(o.i(mo.bonds.empty() ? 0 : mo.bonds[0][k]) + " ");
                    continue;
                }
                o.tok(o.i((long)mo.bonds.size()) + "\n");
                for (int k = 1; k <= nc; k++) {
                    std::string t = o.f(mo.constants[k - 1]) + " ";
                    if (mo.type == BEAD && k % kBeadConst == 0) t += "\n";
                    if (k == nc) t += "\n";
                    o.tok(t);
                }
                for (const auto &b : mo.bonds)
                    for (int k = 0; k < np; k++) o.tok(o.i(b[k]) + (k == np - 1 ? " \n" : " "));
            }
            break;
        case BANANA: o.tok("\n"); break;
        case DELTALXY: o.tok(o.f(s.deltaLXY)); break;
        case REMOVESOLVENT: o.tok(o.f(s.removeSolvent)); break;
        case TEMPSTEPINTERVAL: o.tok(o.f(s.tempStepInterval)); break;
        case SOLVENTGAMMA: o.tok(o.f(s.solventGamma[0])); o.tok(o.f(s.solventGamma[1])); break;
        case GAMMATYPE: for (double g : s.gammaType) o.tok(o.f(g)); break;
        case TENSION: o.tok(o.f(s.tension)); break;
        }
    }
    o.s += '\n';
    return o.s;
}

void write_file(const System &s, const std::string &path)
{
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f.is_open()) throw std::runtime_error("Could not open " + path + " for writing");
    std::string t = write_text(s);
    f.write(t.data(), (std::streamsize)t.size());
    if (!f) throw std::runtime_error("write to " + path + " failed");
}

// ---------------------------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------------------------
namespace {

const double kCutoff = 2.0, kRmin = 1.0;  // generate/include.h:28-29
const int kHead = 2, kTail = 3, kAnchor = 5;  // include/MD.h:21-25

// The reference is built with g++ -O3, which fuses each sin/cos pair of one argument into
// a single glibc sincos() call; sincos and separate sin/cos can differ in the last ulp, so
// the generators call sincos explicitly to reproduce the reference's bytes.
inline void sin_cos(double x, double *s, double *c) { ::sincos(x, s, c); }

// MTRand(seed).rand53() is MT19937 with init_genrand(seed) and the genrand_res53
// construction (MersenneTwister.h:337-341); std::mt19937 implements the same generator.
struct Rand53 {
    std::mt19937 g;
    explicit Rand53(int seed) : g((uint32_t)seed) {}
    double operator()()
    {
        uint32_t a = g() >> 5, b = g() >> 6;
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
    }
};

// laradjiRevaleeFC/PC (include/potentials/laradjiRevalee.h:84-152): row (i,j) uses the
// matrices at i + j*nTypes.
void revalee_tables(System &s, const std::vector<double> &Umax, const std::vector<double> &Umin)
{
    const int n = s.nTypes;
    const double rc = kCutoff, rm = kRmin;
    s.fconst.clear();
    s.uconst.clear();
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int g = i + j * n;
            s.fconst.insert(s.fconst.end(), {rm, (2.0 * (Umax[g] - Umin[g])) / (rm * rm), 0.0, rc,
                                             (6.0 * Umin[g]) / ((rc - rm) * (rc - rm)),
                                             (6.0 * Umin[g]) / ((rc - rm) * (rc - rm) * (rc - rm))});
            s.uconst.insert(s.uconst.end(), {rm, (Umax[g] - Umin[g]) / (rm * rm), Umin[g], rc,
                                             (3.0 * Umin[g]) / ((rc - rm) * (rc - rm)),
                                             (2.0 * Umin[g]) / ((rc - rm) * (rc - rm) * (rc - rm))});
        }
    s.present[TWOBODYFCONST] = s.present[TWOBODYUCONST] = true;
}

void common_header(System &s, int seed, double finalTime, double storeI, double measureI)
{
    s.gamma = 1.0; s.present[GAMMA] = true;
    s.nTypes = 6; s.present[NTYPES] = true;
    s.seed = seed; s.present[SEED] = true;
    s.periodic = true; s.present[PERIODIC] = true;
    s.cutoff = kCutoff; s.present[CUTOFF] = true;
    s.initialTime = 0; s.present[INITIALTIME] = true;
    s.finalTime = finalTime; s.present[FINALTIME] = true;
    s.deltaT = 0.02; s.present[DELTAT] = true;
    s.storeInterval = storeI; s.present[STOREINTERVAL] = true;
    s.measureInterval = measureI; s.present[MEASUREINTERVAL] = true;
    s.initialTemp = 3.0; s.present[INITIALTEMP] = true;
    s.finalTemp = 3.0; s.present[FINALTEMP] = true;
}

void add_particle(System &s, int type, double x, double y, double z, double vx, double vy, double vz)
{
    s.type.push_back(type);
    s.pos.insert(s.pos.end(), {x, y, z});
    s.vel.insert(s.vel.end(), {vx, vy, vz});
    s.nParticles++;
    s.present[NPARTICLES] = s.present[POSITIONS] = s.present[VELOCITIES] = true;
}

void add_molecule(System &s, const Molecule &m)
{
    s.mol.push_back(m);
    s.nMolecules++;
    s.present[NMOLECULES] = s.present[MOLECULE] = true;
}

}  // namespace

System gen_liposome(int seed, int nLipids, double arealDensity)
{
    System s;
    common_header(s, seed, 50000, 1000, 100);  // generate/liposome.cpp:19-37
    std::vector<double> Umax(36, 100.0), Umin(36, 0.0);
    Umin[kTail + kTail * 6] = -6;
    Umax[kTail + kTail * 6] = 200;
    revalee_tables(s, Umax, Umin);
    s.size[0] = s.size[1] = s.size[2] = 400;
    s.present[SIZE] = true;
    const double pos[3] = {200, 200, 200};
    const int lipidLength = 3;
    const double bondLength = 0.7;

    // models/lipidModels.h:1313-1469 with nConstants == 4 (CHAIN)
    Molecule m;
    m.type = CHAIN;
    m.bonds.push_back({s.nParticles, nLipids, lipidLength, 0});
    m.constants = {0.7, 100, 1.0, 100};
    const double radius = std::sqrt(((double)nLipids / arealDensity) / (4.0 * M_PI));
    for (int d = 0; d < 3; d++)
        s.size[d] = s.size[d] < radius + pos[d] + lipidLength * bondLength ? pos[d] * 2.0 : s.size[d];
    int inner = nLipids * std::pow(radius, 2.0) /
                (std::pow(radius, 2.0) + std::pow(radius + (double)(lipidLength * 2 - 1) * bondLength, 2.0));
    int outer = nLipids - inner;
    Rand53 rnd(seed);
    const double Vrms = std::sqrt(3.0 * s.initialTemp);
    s.type.reserve(3 * (size_t)nLipids);
    s.pos.reserve(9 * (size_t)nLipids);
    s.vel.reserve(9 * (size_t)nLipids);
    for (int leaflet = 0; leaflet < 2; leaflet++) {
        const int count = leaflet ? outer : inner;
        double sp = 3.6 / std::sqrt(double(count));
        double length = 0;
        double dz = 2.0 / double(count);
        double z = 1.0 - dz / 2.0;
        for (int i = 0; i < count; i++) {
            double r = std::sqrt(1.0 - z * z);
            double sl, cl;
            sin_cos(length, &sl, &cl);
            for (int j = 0; j < lipidLength; j++) {
                double shell = leaflet ? radius + (bondLength * (double)(lipidLength + j))
                                       : radius + bondLength * (double)j;
                double x = pos[0] + r * cl * shell;
                double y = pos[1] + r * sl * shell;
                double zz = pos[2] + z * shell;
                int type = leaflet ? ((j == lipidLength - 1) ? kHead : kTail) : ((j == 0) ? kHead : kTail);
                double theta = M_PI * rnd();
                double phi = M_PI * 2.0 * rnd();
                double st, ct, sp2, cp;
                sin_cos(theta, &st, &ct);
                sin_cos(phi, &sp2, &cp);
                add_particle(s, type, x, y, zz, Vrms * cp * st, Vrms * sp2 * st, Vrms * ct);
            }
            z = z - dz;
            length = length + sp / r;
        }
    }
    add_molecule(s, m);
    return s;
}

System gen_bilayer(int seed, int nLipids, double arealDensity)
{
    System s;
    common_header(s, seed, 1000, 10, 1);  // generate/bilayer.cpp:25-43
    s.deltaLXY = 0.01;
    s.present[DELTALXY] = true;
    std::vector<double> Umax(36, 100.0), Umin(36, 0.0);
    Umin[kTail + kTail * 6] = -6;
    Umax[kTail + kTail * 6] = 200;
    Umin[kAnchor + kHead * 6] = 0.0;  // UminAnchorHead = 0 for the benchmark inputs
    Umax[kAnchor + kHead * 6] = 200;
    Umin[kHead + kAnchor * 6] = Umin[kAnchor + kHead * 6];
    Umax[kHead + kAnchor * 6] = Umax[kAnchor + kHead * 6];
    revalee_tables(s, Umax, Umin);
    s.size[0] = s.size[1] = 0;
    s.size[2] = 40;
    s.present[SIZE] = true;
    const double pos[3] = {0, 0, s.size[2] / 2.0};
    const double bondLength = 0.7;
    const int lipidLength = 3;
    const double aspectRatio = 1.0;

    // models/lipidModels.h:1747-1924 with nConstants == 4, heads HEAD, tails TAIL
    Molecule m;
    m.type = CHAIN;
    m.constants = {bondLength, 100, 1, 100};
    std::array<int, 4> chain = {s.nParticles, 0, lipidLength, 0};
    double need[3] = {std::sqrt((double)nLipids / arealDensity) * std::sqrt(aspectRatio),
                      std::sqrt((double)nLipids / arealDensity) / std::sqrt(aspectRatio), pos[2] * 2.0};
    for (int d = 0; d < 3; d++) s.size[d] = (s.size[d] < need[d]) ? need[d] : s.size[d];
    Rand53 rnd(seed);
    const double Vrms = std::sqrt(3.0 * s.initialTemp);
    const double ll = std::sqrt(2.0 / arealDensity);
    const int latX = (int)(static_cast<int>(s.size[0]) / ll + 1);
    const int latY = (int)(static_cast<int>(s.size[1]) / ll + 1);
    s.type.reserve(3 * (size_t)nLipids);
    s.pos.reserve(9 * (size_t)nLipids);
    s.vel.reserve(9 * (size_t)nLipids);
    for (int i = 0; i < latX; i++)
        for (int j = 0; j < latY; j++) {
            int lipidIndex = 2 * (j * latX + i);
            for (int leaflet = 0; leaflet < 2; leaflet++) {
                if (lipidIndex + leaflet >= nLipids) continue;
                for (int k = 0; k < lipidLength; k++) {
                    double x = i * ll + 0.001;
                    double y = j * ll + 0.001;
                    double z = leaflet ? pos[2] + bondLength * (lipidLength + k + 1)
                                       : pos[2] + bondLength * (lipidLength - k);
                    int type = (k == lipidLength - 1) ? kHead : kTail;
                    double theta = M_PI * rnd();
                    double phi = M_PI * 2 * rnd();
                    double st, ct, sp, cp;
                    sin_cos(theta, &st, &ct);
                    sin_cos(phi, &sp, &cp);
                    add_particle(s, type, x, y, z, Vrms * cp * st, Vrms * sp * st, Vrms * ct);
                }
                chain[1]++;
            }
        }
    m.bonds.push_back(chain);
    add_molecule(s, m);
    return s;
}

}  // namespace smio
