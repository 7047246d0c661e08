// md_hip: drop-in replacement for the reference `MD name` executable (MD.cpp:72-754) and
// its GPU twin (MDcuda.cu:77-495), running the step through the C ABI of libsoftmold_hip.
//
// Behaviour (the process boundary a user of `MD` relies on):
//  - exactly one argument, else "usage: ..." on stderr and exit 0 (MD.cpp:74-79);
//  - reads name.mpd, falling back to name.tmp.mpd (MDcuda.cu:94-112);
//  - steps int(t0/dt+1e-7) .. int(t1/dt+1e-7) inclusive (MD.cpp:311-335);
//  - t0 == 0: step-0 forces, one measurement line and one frame (MD.cpp:266-273);
//    t0 != 0: forces then Verlet::second, completing the mid-step store (MD.cpp:274-308);
//  - every storeInterval (i != start): after the first half step, writes name.tmp.mpd then
//    name.mpd and appends a frame to frames_name.xyz (MD.cpp:373-381, MDcuda.cu:351-365);
//  - every measureInterval (i != start): appends `time\tvalue` to potential_name.dat and
//    kinetic_name.dat, `time\tLx\tLy\tLz` to size_name.dat, the mean bond length to
//    lBond_name.dat, `time\tcos\tla\tlb` of the bends to bend_name.dat, `time\tT` to
//    temp_name.dat (dataExtraction.h:992-1050, dataCollection.h:677-690) and
//    `time\tseconds` on stderr;
//  - at step 0 and every measurement, the extent of the particles along each axis (largest
//    minus smallest coordinate, from the box and 0) to flicker_name.dat, and the particles'
//    kinetic energies 0.5 |v|^2 into a histogram of width 1e-4 written as
//    kEnergyDensity_name.dat at the end (dataExtraction.h:1456-1487,1512-1519,552-567;
//    MD.cpp builds without FLAT_MEMBRANE and with its unwrapped-position array set);
//  - temperature ramp by tempStepInterval (MD.cpp:320-325,369-371);
//  - MC box move every 8 steps when deltaLXY != 0 (MD.cpp:589-721), resizeHist_name.dat
//    and "Resize acceptance ratio" at exit (MD.cpp:731-748).
// Runs of steps with no store, measure or temperature change go to the device in one
// sm_run call; nothing is copied back for them.
#include "softmold_hip.h"

#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct Fail {
    int code;
};

void check(int rc, const char *what)
{
    if (rc != SM_OK) {
        std::cerr << "md_hip: " << what << ": " << sm_last_error() << std::endl;
        throw Fail{rc};
    }
}

double get(const sm_mpd *m, const char *key)
{
    double v = 0;
    check(sm_mpd_get(m, key, &v), key);
    return v;
}

void append_line(const std::string &file, double t, const double *v, int nv)
{
    std::fstream f(file.c_str(), std::ios::out | std::ios::app);
    f << t;
    for (int k = 0; k < nv; k++) f << '\t' << v[k];
    f << std::endl;
}

// flicker_ line: per axis, the largest coordinate (from 0) minus the smallest (from the box
// length) over all particles (dataExtraction.h:1456-1487)
void append_flicker(const std::string &file, double t, int n, const double *pos, const double *box)
{
    double ext[3];
    for (int d = 0; d < 3; d++) {
        double lo = box[d], hi = 0.0;
        for (int i = 0; i < n; i++) {
            const double p = pos[3 * (size_t)i + d];
            if (p < lo) lo = p;
            if (p > hi) hi = p;
        }
        ext[d] = hi - lo;
    }
    append_line(file, t, ext, 3);
}

// kinetic energy histogram of a measurement (dataExtraction.h:1512-1519): bin
// int(0.5 |v|^2 / 1e-4), grown to the largest bin seen.  A value that is not finite or
// lies past bin 2^28 (0.5 |v|^2 > 26843, thousands of times the thermal energy) is not
// counted: the reference would convert it to int undefined or grow its table to gigabytes.
constexpr double kKeBin = 0.0001;
constexpr double kKeMaxBin = 268435456.0;
void add_kinetic(std::vector<long> &hist, int n, const double *vel)
{
    for (int i = 0; i < n; i++) {
        const double *v = vel + 3 * (size_t)i;
        const double e = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / kKeBin;
        if (!(e >= 0.0 && e < kKeMaxBin)) continue;
        const int bin = (int)e;
        if ((size_t)bin >= hist.size()) hist.resize((size_t)bin + 1, 0);
        hist[bin]++;
    }
}

// kEnergyDensity_ at the end of the run (dataExtraction.h:552-567): every bin's lower edge
// and its share of all counts
void write_kinetic(const std::string &file, const std::vector<long> &hist)
{
    long sum = 0;
    for (long c : hist) sum += c;
    std::fstream f(file.c_str(), std::ios::out);
    for (size_t i = 0; i < hist.size(); i++)
        f << static_cast<float>(i) * kKeBin << '\t' << static_cast<float>(hist[i]) / static_cast<float>(sum) << '\n';
}

// xyzFormat::store (xyzFormat.h:112-143): count, "test", then type x y z per particle.
void append_frame(const std::string &file, int n, const int32_t *type, const double *pos)
{
    std::fstream f(file.c_str(), std::ios::out | std::ios::app);
    f << n << '\n' << "test\n";
    for (int i = 0; i < n; i++)
        f << type[i] << '\t' << pos[3 * i] << '\t' << pos[3 * i + 1] << '\t' << pos[3 * i + 2] << '\n';
}

int run(const char *name)
{
    sm_mpd *mpd = nullptr;
    if (sm_mpd_read(name, &mpd) != SM_OK) {
        const std::string tmp = std::string(name) + ".tmp";
        if (sm_mpd_read(tmp.c_str(), &mpd) != SM_OK) {
            std::cerr << "Cannot open temp file! System load failure! (" << sm_last_error() << ")" << std::endl;
            return -1;
        }
    }
    const std::string nm(name);
    const std::string fFrames = "frames_" + nm + ".xyz", fPot = "potential_" + nm + ".dat",
                      fKin = "kinetic_" + nm + ".dat", fSize = "size_" + nm + ".dat", fTemp = "temp_" + nm + ".dat",
                      fHist = "resizeHist_" + nm + ".dat", fBond = "lBond_" + nm + ".dat",
                      fBend = "bend_" + nm + ".dat", fFlick = "flicker_" + nm + ".dat",
                      fKe = "kEnergyDensity_" + nm + ".dat";
    int rc = 0;
    sm_ctx *ctx = nullptr;
    try {
        sm_system sys;
        if (sm_mpd_system(mpd, &sys) != SM_OK && std::string(sm_last_error()) == "No gamma available!") {
            // MD.cpp:139-143: neither gamma nor gammaType; message on stdout, exit status 0
            std::cout << "Error(main): No gamma available!\n";
            sm_mpd_free(mpd);
            return 0;
        }
        check(sm_mpd_system(mpd, &sys), "system");
        const int n = sys.n_particles;
        const double dt = sys.dt;
        const double t0 = get(mpd, "initialTime"), t1 = get(mpd, "finalTime");
        const int endInt = int(t1 / dt + 0.0000001);
        const int startInt = int(t0 / dt + 0.0000001);
        const int storeint = int(get(mpd, "storeInterval") / dt + 0.0000001);
        const int measureint = int(get(mpd, "measureInterval") / dt + 0.0000001);
        if (storeint <= 0 || measureint <= 0) {
            std::cerr << "md_hip: storeInterval and measureInterval must be at least one time step" << std::endl;
            throw Fail{SM_ERR_INVALID};
        }
        double temp = sys.temperature;
        const double tsi = get(mpd, "tempStepInterval");
        int tempStepInt = 0;
        double tempStep = 0;
        if (tsi > 0) {
            tempStep = tsi * (get(mpd, "finalTemp") - temp) / (t1 - t0);
            const int div = int((t1 - t0) / tsi);
            tempStepInt = div != 0 ? (endInt - startInt) / div : 0;
        }
        std::vector<int32_t> type(sys.type, sys.type + n);
        std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n);
        double box[3] = {sys.box[0], sys.box[1], sys.box[2]};
        std::vector<long> kehist;

        check(sm_create(&sys, &ctx), "create");
        check(sm_init_forces(ctx), "initial forces");
        if (t0 == 0) {
            double m[3], bs[4];
            check(sm_measure(ctx, m), "measure");
            check(sm_bond_stats(ctx, bs), "bond statistics");
            const double u = m[0] + m[1];
            append_line(fPot, t0, &u, 1);
            append_line(fSize, t0, box, 3);
            append_line(fBond, t0, bs, 1);
            append_line(fBend, t0, bs + 1, 3);
            append_line(fTemp, t0, &temp, 1);
            append_line(fKin, t0, &m[2], 1);
            append_flicker(fFlick, t0, n, sys.pos, box);
            add_kinetic(kehist, n, sys.vel);
            append_frame(fFrames, n, type.data(), sys.pos);
        } else {
            check(sm_second_half(ctx), "second half step");
        }

        std::cerr << "starting main loop: \n";
        time_t current = time(nullptr);
        int i = startInt;
        while (i <= endInt) {
            // plain steps up to the next event go to the device in one call
            int j = i;
            auto event = [&](int s) {
                const bool tstep = tempStepInt != 0 && s % tempStepInt == 0 && s < endInt;
                return tstep || (s % storeint == 0 && s != startInt) || (s % measureint == 0 && s != startInt);
            };
            while (j <= endInt && !event(j)) j++;
            if (j > i) {
                check(sm_run(ctx, i, j - i), "run");
                i = j;
                continue;
            }
            const double t = (double)i * dt;
            if (tempStepInt != 0 && i % tempStepInt == 0 && i < endInt) {
                temp += tempStep;
                check(sm_set_temperature(ctx, temp), "temperature");
            }
            check(sm_step_first(ctx, i), "step");
            if (i % storeint == 0 && i != startInt) {
                check(sm_download_state(ctx, pos.data(), vel.data(), nullptr, box), "download");
                check(sm_mpd_set_state(mpd, pos.data(), vel.data(), box), "state");
                check(sm_mpd_set(mpd, "initialTime", t), "initialTime");
                check(sm_mpd_set(mpd, "initialTemp", temp), "initialTemp");
                check(sm_mpd_write(mpd, (nm + ".tmp").c_str()), "write");
                check(sm_mpd_write(mpd, name), "write");
                append_frame(fFrames, n, type.data(), pos.data());
            }
            const bool measure = i % measureint == 0 && i != startInt;
            // the box is measured before this step's MC move (MD.cpp:513-589) and no
            // earlier part of the step changes it
            double m[3], b[3], bs[4];
            if (measure) {
                // positions before this step's MC move, as the measurement sees them
                check(sm_download_state(ctx, pos.data(), nullptr, nullptr, b), "positions");
                check(sm_bond_stats(ctx, bs), "bond statistics");   // positions as measured
            }
            check(sm_step_finish(ctx, i, measure ? m : nullptr), "step");
            if (measure) {
                // velocities after the kick (the MC move leaves them)
                check(sm_download_state(ctx, nullptr, vel.data(), nullptr, nullptr), "velocities");
                const time_t last = current;
                current = time(nullptr);
                std::cerr << t << '\t' << current - last << std::endl;
                const double u = m[0] + m[1];
                append_line(fPot, t, &u, 1);
                append_line(fSize, t, b, 3);
                append_line(fBond, t, bs, 1);
                append_line(fBend, t, bs + 1, 3);
                append_line(fTemp, t, &temp, 1);
                append_line(fKin, t, &m[2], 1);
                append_flicker(fFlick, t, n, pos.data(), b);
                add_kinetic(kehist, n, vel.data());
                current = time(nullptr);
            }
            i++;
        }

        int64_t trials = 0, accepted = 0;
        check(sm_mc_stats(ctx, &trials, &accepted), "mc stats");
        if (sys.delta_lxy != 0) {
            std::vector<double> ha(256), hr(256);
            int32_t nb = 0;
            check(sm_mc_hist(ctx, 0, 256, ha.data(), hr.data(), &nb), "mc histogram");
            const double interval = sys.delta_lxy / 101.0;
            std::fstream f(fHist.c_str(), std::ios::out);
            for (int k = 0; k < nb; k++)
                f << (static_cast<double>(k) * interval) - sys.delta_lxy << '\t' << ha[k] << '\t' << hr[k] << std::endl;
        }
        const double tr = (double)trials, ac = (double)accepted;
        std::cerr << "Resize acceptance ratio: " << ac / tr << std::endl;
        write_kinetic(fKe, kehist);
    } catch (const Fail &f) {
        rc = f.code;
    }
    if (ctx) sm_destroy(ctx);
    sm_mpd_free(mpd);
    return rc;
}

}  // namespace

int main(int argc, char *argv[])
{
    if (argc != 2) {
        std::cerr << "usage: " << argv[0] << " name\n";
        return 0;
    }
    return run(argv[1]);
}
