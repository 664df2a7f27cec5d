// cgfd_conf.h -- the C++ parameter-file front end: the inputs of one run
// from the reference's own parameter files (SeisFD3D.conf and the grid,
// absorb, source, station and media files it names), as the reference's
// master prepares them (simulation.cc:759-790, gridmesh.cpp, absorb.cpp,
// source.cpp, seisplot.cpp, mediaparas.cpp), for a host that hands them to
// the engine's C-ABI.
//
// Arithmetic follows the reference's Real = float build: parsed values are
// (float)atof / strtof, expressions keep C's float/double promotions, and
// float-argument sin/cos/exp/pow/log bind sinf/cosf/expf/powf/logf.
// Arrays use the reference layout [x][y][z], z fastest, 3 ghosts per side.
#ifndef CGFD_CONF_H
#define CGFD_CONF_H

#include <stdexcept>
#include <string>
#include <vector>

namespace cgfd_conf {

// A parameter file is missing or malformed, or asks for something the front
// end refuses (the message says which).
struct ConfError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Scenario {
    int ni = 0, nj = 0, nk = 0, nt = 0;
    float dh = 0, dt = 0;
    int pgv = 0;
    int hybrid = 0, conindex = 3;            // ConIndex with ghosts (3 = all curvilinear)
    std::vector<float> x, y, z;              // coordinates with ghosts
    std::vector<float> metrics[10];          // xix xiy xiz etax etay etaz zetax zetay zetaz jac
    std::vector<float> alpha, beta, rho;
    int nabs[6] = {0, 0, 0, 0, 0, 0};
    std::vector<float> pml[9];               // APDx APDy APDz Bx By Bz DBx DBy DBz
    std::vector<int> force_loc, moment_loc;  // [n][3], ghost-inclusive
    std::vector<float> force_vec, force_stf; // [n][3], [n][2(nt+1)]
    std::vector<float> moment_vec, moment_stf;
    std::vector<int> focal_loc;              // [n][3]
    std::vector<float> focal_series;         // [6][n][focal_nt]
    int focal_nt = 0;
    float focal_dt = 0;
    std::vector<int> recv_loc;               // [nrecv][3]
    std::vector<float> recv_pos;             // [nrecv][3], metres
    std::vector<int> recv_pnum, recv_lnum;
    int seismo_tinv = 1;
    std::vector<int> snapshots;              // [nsnap][11]

    size_t cells() const { return size_t(ni + 6) * (nj + 6) * (nk + 6); }
    int nforce() const { return int(force_loc.size() / 3); }
    int nmoment() const { return int(moment_loc.size() / 3); }
    int nfocal() const { return int(focal_loc.size() / 3); }
    int nrecv() const { return int(recv_loc.size() / 3); }
    int nsnap() const { return int(snapshots.size() / 11); }
};

// Loads a SeisFD3D.conf.  `seispath`, when not empty, replaces the file's own
// seispath (the directory of the other files).  `zwmedia` reads a volume
// file's depth coordinate as "depth" (the reference's Makefile builds with
// ZWmedia) instead of "z".  Throws ConfError.
Scenario load(const std::string &conf, const std::string &seispath = "", bool zwmedia = true);

} // namespace cgfd_conf

#endif
