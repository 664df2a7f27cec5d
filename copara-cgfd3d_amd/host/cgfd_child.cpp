// cgfd_child -- one child-rank run of the reference's simulation driven
// through the C-ABI from C++ (the loop of simulation.cc:139-717 for one
// process and one GPU, with ChildProcs replaced by cgfd_*).
//
// The master's parameter arrays (mpiRecvPars / mpiRecvABC / mpiRecvPP /
// mpiRecvSP / mpiFDRecv, simulation.cc:155-186) are read from raw files in
// a directory, in the reference layout ([x][y][z], z fastest, 3 ghosts per
// side) and the run's Real type (float, or double for precision 1):
//   dims.txt      ni nj nk dh dt nt precision nforce nmoment nrecv nsnap pgv
//   metrics.bin   xix xiy xiz etax etay etaz zetax zetay zetaz jac (NX*NY*NZ each)
//   media.bin     alpha beta rho
//   nabs.bin      6 int32
//   pml.bin       APDx APDy APDz Bx By Bz DBx DBy DBz (lengths NX NY NZ, repeated)
//   force_loc.bin / force_vec.bin / force_stf.bin     int32[n][3] / Real[n][3] / Real[n][2(nt+1)]
//   moment_loc.bin / moment_vec.bin / moment_stf.bin  int32[n][3] / Real[n][6] / Real[n][2(nt+1)]
//   recv.bin      int32[nrecv][3]
//   snap.bin      int32[nsnap][11] (snap_NNN lines of the station file)
// Outputs into the same directory: traces.bin (Real[9][nrecv][nt]),
// pgv.bin (Real[3][NX][NY]) and the reference's seismo.nc, snap_NNN.nc and
// peakvel.nc.
//
// usage: cgfd_child <dir> [steps]
#include "cgfd3d.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

std::string g_dir;

[[noreturn]] void die(const std::string &msg) {
    std::fprintf(stderr, "cgfd_child: %s\n", msg.c_str());
    std::exit(1);
}

void check(cgfd_ctx *c, int rc, const char *what) {
    if (rc != CGFD_OK) die(std::string(what) + ": " + (c ? cgfd_last_error(c) : "no context"));
}

template <typename T>
std::vector<T> read_bin(const char *name, size_t count) {
    std::vector<T> v(count);
    if (count == 0) return v;
    std::string path = g_dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die("cannot open " + path);
    size_t got = std::fread(v.data(), sizeof(T), count, f);
    std::fclose(f);
    if (got != count) die(path + ": expected " + std::to_string(count) + " values, found " + std::to_string(got));
    return v;
}

template <typename T>
void write_bin(const char *name, const std::vector<T> &v) {
    std::string path = g_dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) die("cannot write " + path);
    std::fclose(f);
}

template <typename Real>
int run(const cgfd_dims &d, int nforce, int nmoment, int nrecv, int nsnap, int pgv, int steps) {
    const size_t NX = d.ni + 6, NY = d.nj + 6, NZ = d.nk + 6, N = NX * NY * NZ;
    const int nstf = 2 * (d.nt + 1);
    cgfd_ctx *c = nullptr;
    check(c, cgfd_create(&d, sizeof(Real) == 8 ? CGFD_FP64 : CGFD_FP32, 0, &c), "create");

    auto met = read_bin<Real>("metrics.bin", 10 * N);               // ParH2D derivF
    const void *drv[10];
    for (int i = 0; i < 10; ++i) drv[i] = met.data() + i * N;
    check(c, cgfd_set_metrics(c, drv), "metrics");
    auto med = read_bin<Real>("media.bin", 3 * N);                  // ParH2D mdparF
    check(c, cgfd_set_media(c, med.data(), med.data() + N, med.data() + 2 * N), "media");

    auto nabs = read_bin<int>("nabs.bin", 6);                       // ParH2D apara
    auto pml = read_bin<Real>("pml.bin", 3 * (NX + NY + NZ));
    const size_t len[3] = {NX, NY, NZ};
    const void *tabs[9];
    size_t off = 0;
    for (int t = 0; t < 9; ++t) {
        tabs[t] = pml.data() + off;
        off += len[t % 3];
    }
    check(c, cgfd_set_pml(c, nabs.data(), tabs), "pml");

    if (nforce) {                                                   // forceF
        auto loc = read_bin<int>("force_loc.bin", 3 * nforce);
        auto vec = read_bin<Real>("force_vec.bin", 3 * nforce);
        auto stf = read_bin<Real>("force_stf.bin", size_t(nforce) * nstf);
        check(c, cgfd_set_forces(c, nforce, loc.data(), vec.data(), stf.data(), nstf), "forces");
    }
    if (nmoment) {                                                  // momentF
        auto loc = read_bin<int>("moment_loc.bin", 3 * nmoment);
        auto vec = read_bin<Real>("moment_vec.bin", 6 * nmoment);
        auto stf = read_bin<Real>("moment_stf.bin", size_t(nmoment) * nstf);
        check(c, cgfd_set_moments(c, nmoment, loc.data(), vec.data(), stf.data(), nstf), "moments");
    }
    auto recv = read_bin<int>("recv.bin", 3 * nrecv);               // C2DPointPick
    if (nrecv) check(c, cgfd_set_receivers(c, nrecv, recv.data()), "receivers");
    auto snap = read_bin<int>("snap.bin", 11 * nsnap);              // C2DSnapPick
    if (nsnap) check(c, cgfd_set_snapshots(c, nsnap, snap.data()), "snapshots");
    check(c, cgfd_set_pgv(c, pgv), "pgv");
    check(c, cgfd_init(c), "init");                                 // VelCoeff

    check(c, cgfd_run(c, steps), "run");                            // the time loop
    check(c, cgfd_sync(c), "sync");

    std::vector<Real> traces(size_t(9) * nrecv * d.nt);             // PWgather
    if (nrecv) check(c, cgfd_get_seismo(c, traces.data()), "get_seismo");
    write_bin("traces.bin", traces);
    std::vector<Real> pv(3 * NX * NY);                              // SynPV
    check(c, cgfd_get_pgv(c, pv.data()), "get_pgv");
    write_bin("pgv.bin", pv);
    if (nrecv) check(c, cgfd_write_seismo(c, (g_dir + "/seismo.nc").c_str(), 1, nullptr, nullptr, nullptr, nullptr,
                                          nullptr), "seismo.nc");
    for (int i = 0; i < nsnap; ++i) {
        char name[32];
        std::snprintf(name, sizeof name, "/snap_%03d.nc", i + 1);
        check(c, cgfd_write_snap(c, i, (g_dir + name).c_str()), "snap_NNN.nc");
    }
    if (pgv) check(c, cgfd_write_peakvel(c, (g_dir + "/peakvel.nc").c_str(), 1), "peakvel.nc");
    cgfd_destroy(c);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) die("usage: cgfd_child <dir> [steps]");
    g_dir = argv[1];
    FILE *f = std::fopen((g_dir + "/dims.txt").c_str(), "r");
    if (!f) die("cannot open " + g_dir + "/dims.txt");
    cgfd_dims d{};
    int prec = 0, nforce = 0, nmoment = 0, nrecv = 0, nsnap = 0, pgv = 1;
    int n = std::fscanf(f, "%d %d %d %lf %lf %d %d %d %d %d %d %d", &d.ni, &d.nj, &d.nk, &d.dh, &d.dt, &d.nt, &prec,
                        &nforce, &nmoment, &nrecv, &nsnap, &pgv);
    std::fclose(f);
    if (n != 12) die("dims.txt: expected 12 values");
    const int steps = argc > 2 ? std::atoi(argv[2]) : d.nt;
    return prec == CGFD_FP64 ? run<double>(d, nforce, nmoment, nrecv, nsnap, pgv, steps)
                             : run<float>(d, nforce, nmoment, nrecv, nsnap, pgv, steps);
}
