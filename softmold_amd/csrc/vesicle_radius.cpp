// vesicle_radius: drop-in for the reference analysis program `vesicleRadius name`
// (analysis/vesicleRadius.cpp:7-205), reading the files md_hip writes.
//
//  - exactly one argument, else a usage note on stderr and exit 0 (:10-16);
//  - reads name.mpd for the particle count, box and storeInterval (:21-33,40);
//  - reads frames_name.xyz frame by frame (xyzFormat::load, xyzFormat.h:146-178);
//  - takes the box of each frame from size_name.dat when it exists (:43-61,185-193): the
//    first line with time >= 0 gives the first frame's time and box, and each later frame
//    (time += storeInterval) takes the first line at or after its time;
//  - per frame prints `time radius` on stdout and `Read frame k with time t tau.` on
//    stderr (:168-172), then the load rate (:198-200).
// The radius is the mean distance of the HEAD beads (type 2, MD.h:21) from their centre of
// mass, with the reference's image rules (:71-165): the centre-of-mass loop moves a head by
// one box length towards particle 0 when it is more than half a box away; the radius loop
// ADDS its image term to the difference (d += minImg), which matches a true minimum image
// only while no head is more than half a box from the centre, and is kept as is.
// The same computation on device-resident state is sm_vesicle_radius.
#include "softmold_hip.h"

#include <cmath>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct Frame {
    std::vector<int> type;
    std::vector<double> x, y, z;
};

// xyzFormat::load: count, a name token, then `type x y z` per particle; false at the end
// of the file (a frame that ends exactly at end-of-file without a newline is dropped, as
// the reference's eof test does).  A token that does not parse also ends the loop; the
// reference would read on forever there, since its loop only tests for end-of-file.
bool load_frame(std::istream &in, Frame &f, int n)
{
    int count = 0;
    std::string name;
    in >> count >> name;
    for (int i = 0; i < n; i++) in >> f.type[i] >> f.x[i] >> f.y[i] >> f.z[i];
    return !in.eof() && !in.fail();
}

double image(double d, double L)
{
    double m = 0.0;
    if (d > L / 2.0) m = L;
    if (d < -L / 2.0) m = -L;
    return m;
}

double radius_of(const Frame &f, int n, const double size[3], int head)
{
    double cx = 0, cy = 0, cz = 0, nh = 0;
    for (int i = 0; i < n; i++) {
        if (f.type[i] != head) continue;
        cx += f.x[i] - image(f.x[i] - f.x[0], size[0]);
        cy += f.y[i] - image(f.y[i] - f.y[0], size[1]);
        cz += f.z[i] - image(f.z[i] - f.z[0], size[2]);
        nh++;
    }
    cx /= nh;
    cy /= nh;
    cz /= nh;
    double radius = 0;
    for (int i = 0; i < n; i++) {
        if (f.type[i] != head) continue;
        double dx = f.x[i] - cx, dy = f.y[i] - cy, dz = f.z[i] - cz;
        dx += image(dx, size[0]);
        dy += image(dy, size[1]);
        dz += image(dz, size[2]);
        radius += std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return radius / nh;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::cerr << "Calculates radius of vesicle and frames per second loading\n";
        std::cerr << "\tusage: " << argv[0] << " name >> radius.dat\n";
        return 0;
    }
    const std::string name = argv[1];
    sm_mpd *m = nullptr;
    if (sm_mpd_read(name.c_str(), &m) != SM_OK) {
        std::cerr << "vesicle_radius: " << sm_last_error() << std::endl;
        return 1;
    }
    sm_system s{};
    double store = 0;
    if (sm_mpd_system(m, &s) != SM_OK || sm_mpd_get(m, "storeInterval", &store) != SM_OK) {
        std::cerr << "vesicle_radius: " << sm_last_error() << std::endl;
        sm_mpd_free(m);
        return 1;
    }
    const int n = s.n_particles;
    double size[3] = {s.box[0], s.box[1], s.box[2]};
    sm_mpd_free(m);

    const std::string frames = "frames_" + name + ".xyz";
    std::ifstream xyz(frames.c_str());
    if (!xyz.is_open()) {
        std::cerr << frames << " not found!\n";
        return 1;
    }
    double time = 0;
    std::ifstream sizeFile(("size_" + name + ".dat").c_str());
    if (sizeFile.is_open()) {
        double t = -1, c[3] = {size[0], size[1], size[2]};
        while (t < 0 && !sizeFile.eof()) sizeFile >> t >> c[0] >> c[1] >> c[2];
        time = t;
        for (int k = 0; k < 3; k++) size[k] = c[k];
    }

    const int head = 2;
    Frame f;
    f.type.assign(n, 0);
    f.x.assign(n, 0.0);
    f.y.assign(n, 0.0);
    f.z.assign(n, 0.0);
    const std::time_t start = std::time(nullptr);
    int frame = 0;
    while (load_frame(xyz, f, n)) {
        std::cout << time << ' ' << radius_of(f, n, size, head) << std::endl;
        std::cerr << "Read frame " << frame << " with time " << time << " tau." << std::endl;
        time += store;
        if (sizeFile.is_open()) {
            double t = 0, c[3] = {size[0], size[1], size[2]};
            while (t < time && !sizeFile.eof()) sizeFile >> t >> c[0] >> c[1] >> c[2];
            for (int k = 0; k < 3; k++) size[k] = c[k];
        }
        frame++;
    }
    const std::time_t end = std::time(nullptr);
    std::cerr << "Loads at " << static_cast<double>(frame) / static_cast<double>(end - start)
              << " frames per second ";
    std::cerr << " in " << end - start << " seconds!\n";
    return 0;
}
