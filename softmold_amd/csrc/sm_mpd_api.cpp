// C ABI of the host-side `.mpd` system files and the two reference generators: the
// part of include/softmold_hip.h that needs no GPU.
#include "../../include/softmold_hip.h"
#include "mpd_io.hpp"
#include "sm_errors.hpp"

#include <cstring>
#include <functional>
#include <string>
#include <vector>

using namespace smapi;

struct sm_mpd {
    smio::System sys;
    std::vector<int32_t> chain, bond, bend, ball;
    std::vector<double> chain_const, bond_const, bend_const, ball_const;
};

namespace {

const struct { const char *name; smio::Cmd cmd; } kScalars[] = {
    {"gamma", smio::GAMMA}, {"initialTemp", smio::INITIALTEMP}, {"finalTemp", smio::FINALTEMP},
    {"seed", smio::SEED}, {"cutoff", smio::CUTOFF}, {"initialTime", smio::INITIALTIME},
    {"finalTime", smio::FINALTIME}, {"deltaT", smio::DELTAT}, {"storeInterval", smio::STOREINTERVAL},
    {"measureInterval", smio::MEASUREINTERVAL}, {"deltaLXY", smio::DELTALXY},
    {"removeSolvent", smio::REMOVESOLVENT}, {"tempStepInterval", smio::TEMPSTEPINTERVAL},
    {"tension", smio::TENSION}, {"nTypes", smio::NTYPES}, {"nParticles", smio::NPARTICLES},
    {"nMolecules", smio::NMOLECULES}};

double *scalar_ref(smio::System &s, smio::Cmd c)
{
    switch (c) {
    case smio::GAMMA: return &s.gamma;
    case smio::INITIALTEMP: return &s.initialTemp;
    case smio::FINALTEMP: return &s.finalTemp;
    case smio::CUTOFF: return &s.cutoff;
    case smio::INITIALTIME: return &s.initialTime;
    case smio::FINALTIME: return &s.finalTime;
    case smio::DELTAT: return &s.deltaT;
    case smio::STOREINTERVAL: return &s.storeInterval;
    case smio::MEASUREINTERVAL: return &s.measureInterval;
    case smio::DELTALXY: return &s.deltaLXY;
    case smio::REMOVESOLVENT: return &s.removeSolvent;
    case smio::TEMPSTEPINTERVAL: return &s.tempStepInterval;
    case smio::TENSION: return &s.tension;
    default: return nullptr;
    }
}

int io_guard(const std::function<void()> &f)
{
    try {
        f();
        return SM_OK;
    } catch (const ApiError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SM_ERR_IO;
    }
}

}  // namespace

extern "C" {

int sm_mpd_read(const char *name, sm_mpd **out)
{
    if (!name || !out) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    *out = nullptr;
    return io_guard([&] {
        auto *m = new sm_mpd();
        try {
            m->sys = smio::read_file(std::string(name) + ".mpd");
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int sm_mpd_write(const sm_mpd *m, const char *name)
{
    if (!m || !name) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    return io_guard([&] { smio::write_file(m->sys, std::string(name) + ".mpd"); });
}

int sm_mpd_free(sm_mpd *m)
{
    delete m;
    return SM_OK;
}

int sm_mpd_system(const sm_mpd *mc, sm_system *out)
{
    if (!mc || !out) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    sm_mpd *m = const_cast<sm_mpd *>(mc);
    return guarded([&] {
        const smio::System &s = m->sys;
        m->chain.clear();
        m->chain_const.clear();
        m->bond.clear();
        m->bond_const.clear();
        m->bend.clear();
        m->bend_const.clear();
        m->ball.clear();
        m->ball_const.clear();
        for (const auto &mo : s.mol) {
            if (mo.type == smio::CHAIN) {
                for (const auto &b : mo.bonds) {
                    m->chain.insert(m->chain.end(), {b[0], b[1], b[2]});
                    m->chain_const.insert(m->chain_const.end(), mo.constants.begin(), mo.constants.end());
                }
            } else if (mo.type == smio::BOND || mo.type == smio::BEND) {
                const bool bond = mo.type == smio::BOND;
                for (const auto &b : mo.bonds) {
                    auto &idx = bond ? m->bond : m->bend;
                    auto &cst = bond ? m->bond_const : m->bend_const;
                    idx.insert(idx.end(), {b[0], b[1]});
                    if (!bond) idx.push_back(b[2]);
                    cst.insert(cst.end(), mo.constants.begin(), mo.constants.begin() + 2);
                }
            } else if (mo.type == smio::BALL) {
                // doBallForce takes the centre from the first record (system.h:1944) and
                // doBallPotential from each record (:2760); they agree when all records
                // share it, the only layout accepted here
                if (mo.bonds.empty()) continue;
                const int centre = mo.bonds[0][0];
                for (const auto &b : mo.bonds) {
                    if (b[0] != centre) unsupported("BALL records with different centre particles");
                    m->ball.insert(m->ball.end(), {centre, b[1]});
                    m->ball_const.insert(m->ball_const.end(), mo.constants.begin(), mo.constants.begin() + 2);
                }
            } else {
                unsupported("molecule type " + std::to_string(mo.type) +
                            " is not implemented (CHAIN, BOND, BEND and BALL are)");
            }
        }
        // MD.cpp:129-143: a positive gamma, else the gammaType thermostat, which applies
        // gammaType[0] to every particle (langevin.h:254 reads type 0) with sigma fixed at
        // the first temperature; with neither, MD stops
        const bool typed = s.gamma <= 0 && s.present[smio::GAMMATYPE] && !s.gammaType.empty();
        if (s.gamma <= 0 && !typed) unsupported("No gamma available!");
        std::memset(out, 0, sizeof *out);
        out->n_particles = s.nParticles;
        out->n_types = s.nTypes;
        out->pos = s.pos.data();
        out->type = s.type.data();
        out->vel = s.vel.data();
        out->acc = nullptr;
        out->n_chain_blocks = (int32_t)(m->chain.size() / 3);
        out->chain = m->chain.data();
        out->chain_const = m->chain_const.data();
        out->fconst = s.fconst.data();
        out->uconst = s.uconst.data();
        for (int d = 0; d < 3; d++) out->box[d] = s.size[d];
        out->cutoff = s.cutoff;
        out->dt = s.deltaT;
        out->gamma = typed ? s.gammaType[0] : s.gamma;
        out->sigma_temperature = typed ? s.initialTemp : 0.0;
        out->temperature = s.initialTemp;
        out->delta_lxy = s.deltaLXY;
        out->tension = s.tension;
        out->seed = (uint64_t)(uint32_t)s.seed;
        out->n_bonds = (int32_t)(m->bond.size() / 2);
        out->bond = m->bond.data();
        out->bond_const = m->bond_const.data();
        out->n_bends = (int32_t)(m->bend.size() / 3);
        out->bend = m->bend.data();
        out->bend_const = m->bend_const.data();
        out->n_balls = (int32_t)(m->ball.size() / 2);
        out->ball = m->ball.data();
        out->ball_const = m->ball_const.data();
    });
}

int sm_mpd_get(const sm_mpd *m, const char *key, double *value)
{
    if (!m || !key || !value) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    for (const auto &k : kScalars) {
        if (std::strcmp(k.name, key) != 0) continue;
        smio::System &s = const_cast<sm_mpd *>(m)->sys;
        switch (k.cmd) {
        case smio::SEED: *value = s.seed; return SM_OK;
        case smio::NTYPES: *value = s.nTypes; return SM_OK;
        case smio::NPARTICLES: *value = s.nParticles; return SM_OK;
        case smio::NMOLECULES: *value = s.nMolecules; return SM_OK;
        default: *value = *scalar_ref(s, k.cmd); return SM_OK;
        }
    }
    g_err = std::string("unknown scalar key ") + key;
    return SM_ERR_INVALID;
}

int sm_mpd_set(sm_mpd *m, const char *key, double value)
{
    if (!m || !key) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    for (const auto &k : kScalars) {
        if (std::strcmp(k.name, key) != 0) continue;
        double *p = scalar_ref(m->sys, k.cmd);
        if (k.cmd == smio::SEED) m->sys.seed = (int)value;
        else if (!p) {
            g_err = std::string(key) + " is not settable";
            return SM_ERR_INVALID;
        } else *p = value;
        // the Blob setters (system.h:4461-4546) drop these keys from the file at zero
        if (k.cmd == smio::TENSION || k.cmd == smio::DELTALXY) m->sys.present[k.cmd] = value != 0.0;
        else if (k.cmd == smio::DELTAT) m->sys.present[k.cmd] = value > 0.0;
        else m->sys.present[k.cmd] = true;
        return SM_OK;
    }
    g_err = std::string("unknown scalar key ") + key;
    return SM_ERR_INVALID;
}

int sm_mpd_set_state(sm_mpd *m, const double *pos, const double *vel, const double box[3])
{
    if (!m) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    const size_t n3 = 3 * (size_t)m->sys.nParticles;
    if (pos) std::memcpy(m->sys.pos.data(), pos, n3 * sizeof(double));
    if (vel) std::memcpy(m->sys.vel.data(), vel, n3 * sizeof(double));
    if (box) for (int d = 0; d < 3; d++) m->sys.size[d] = box[d];
    return SM_OK;
}

int sm_gen_liposome(int32_t seed, int32_t n_lipids, double rho, sm_mpd **out)
{
    if (!out || n_lipids <= 0 || !(rho > 0)) {
        g_err = "bad argument";
        return SM_ERR_INVALID;
    }
    return io_guard([&] {
        auto *m = new sm_mpd();
        m->sys = smio::gen_liposome(seed, n_lipids, rho);
        *out = m;
    });
}

int sm_gen_bilayer(int32_t seed, int32_t n_lipids, double rho, sm_mpd **out)
{
    if (!out || n_lipids <= 0 || !(rho > 0)) {
        g_err = "bad argument";
        return SM_ERR_INVALID;
    }
    return io_guard([&] {
        auto *m = new sm_mpd();
        m->sys = smio::gen_bilayer(seed, n_lipids, rho);
        *out = m;
    });
}

}  // extern "C"
