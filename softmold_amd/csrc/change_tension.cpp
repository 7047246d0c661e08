// change_tension: drop-in for the reference editor `changeTension name tension`
// (analysis/changeTension.cpp:14-47): reads name.mpd, sets the tension key and writes
// name.mpd back (every other key, and the particle data, unchanged).
//  - exactly two arguments, else a usage note on stderr and exit 0 (:16-21);
//  - the value is parsed with atof, as the reference does (:23).
#include "softmold_hip.h"

#include <cstdlib>
#include <iostream>

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::cerr << argv[0] << " name tension\n";
        std::cerr << "tension>0 stretches, tension<0 compresses\n";
        return 0;
    }
    const double tension = std::atof(argv[2]);
    sm_mpd *m = nullptr;
    int rc = sm_mpd_read(argv[1], &m);
    if (rc == SM_OK) rc = sm_mpd_set(m, "tension", tension);
    if (rc == SM_OK) rc = sm_mpd_write(m, argv[1]);
    if (rc != SM_OK) std::cerr << "change_tension: " << sm_last_error() << std::endl;
    if (m) sm_mpd_free(m);
    return rc == SM_OK ? 0 : 1;
}
