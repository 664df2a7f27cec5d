// cgfd_child -- one child-rank run of the reference's simulation driven
// through the C-ABI from C++ (the loop of simulation.cc:139-717 for one
// process and one GPU, with ChildProcs replaced by cgfd_*).
//
// Two ways in:
//
// 1. From the reference's own parameter files, as its master prepares them
//    (cgfd_conf.cpp: SeisFD3D.conf and the grid, absorb, source, station
//    and media files it names):
//      cgfd_child --conf <SeisFD3D.conf> [--seispath <dir>] [--simplecar] --out <dir> [--steps N]
//    runs the simulation and writes the outputs into <dir> (--simplecar: as
//    the reference's SimpleCar build computes it);
//      cgfd_child --conf <SeisFD3D.conf> [--seispath <dir>] --dump <dir>
//    writes the prepared arrays in the raw format below (and coords.bin,
//    recv_pos.bin / recv_pnum.bin / recv_lnum.bin) without touching a GPU.
//
// 2. From the master's parameter arrays (mpiRecvPars / mpiRecvABC /
//    mpiRecvPP / mpiRecvSP / mpiFDRecv, simulation.cc:155-186) in raw files
//    in a directory, in the reference layout ([x][y][z], z fastest, 3 ghosts
//    per side) and the run's Real type (float, or double for precision 1):
//      dims.txt      ni nj nk dh dt nt precision nforce nmoment nrecv nsnap pgv
//                    [nfocal focal_nt focal_dt conindex seismo_tinv]
//      metrics.bin   xix xiy xiz etax etay etaz zetax zetay zetaz jac (NX*NY*NZ each)
//      media.bin     alpha beta rho
//      nabs.bin      6 int32
//      pml.bin       APDx APDy APDz Bx By Bz DBx DBy DBz (lengths NX NY NZ, repeated)
//      force_loc.bin / force_vec.bin / force_stf.bin     int32[n][3] / Real[n][3] / Real[n][2(nt+1)]
//      moment_loc.bin / moment_vec.bin / moment_stf.bin  int32[n][3] / Real[n][6] / Real[n][2(nt+1)]
//      focal_loc.bin / focal_series.bin                  int32[n][3] / Real[6][n][focal_nt]
//      recv.bin      int32[nrecv][3]
//      snap.bin      int32[nsnap][11] (snap_NNN lines of the station file)
//    conindex > 0 selects the hybrid grid (cgfd_set_hybrid).
//      cgfd_child <dir> [steps]
//
// Outputs: traces.bin (Real[9][nrecv][nt]), pgv.bin (Real[3][NX][NY]) and the
// reference's seismo.nc, snap_NNN.nc and peakvel.nc.
#include "cgfd3d.h"
#include "cgfd_conf.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(const std::string &msg) {
    std::fprintf(stderr, "cgfd_child: %s\n", msg.c_str());
    std::exit(1);
}

void check(cgfd_ctx *c, int rc, const char *what) {
    if (rc != CGFD_OK) die(std::string(what) + ": " + (c ? cgfd_last_error(c) : "no context"));
}

template <typename T>
std::vector<T> read_bin(const std::string &dir, const char *name, size_t count) {
    std::vector<T> v(count);
    if (count == 0) return v;
    std::string path = dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) die("cannot open " + path);
    size_t got = std::fread(v.data(), sizeof(T), count, f);
    std::fclose(f);
    if (got != count) die(path + ": expected " + std::to_string(count) + " values, found " + std::to_string(got));
    return v;
}

template <typename T>
void write_bin(const std::string &dir, const char *name, const std::vector<T> &v) {
    std::string path = dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) die("cannot write " + path);
    std::fclose(f);
}

// Everything the child hands to the engine.
template <typename Real>
struct Inputs {
    cgfd_dims d{};
    int pgv = 1, conindex = 0, focal_nt = 0, seismo_tinv = 1, simplecar = 0;
    double focal_dt = 0;
    std::vector<Real> metrics, media, pml, force_vec, force_stf, moment_vec, moment_stf, focal_series;
    std::vector<int> nabs, force_loc, moment_loc, focal_loc, recv, snap;
    // receiver metadata of seismo.nc (the front-end path only)
    std::vector<int> pnum, lnum;
    std::vector<Real> posx, posy, posz;
};

template <typename Real>
int run(const Inputs<Real> &in, int steps, const std::string &out) {
    const cgfd_dims &d = in.d;
    const size_t NX = d.ni + 6, NY = d.nj + 6, NZ = d.nk + 6, N = NX * NY * NZ;
    const int nstf = 2 * (d.nt + 1);
    const int nforce = int(in.force_loc.size() / 3), nmoment = int(in.moment_loc.size() / 3);
    const int nfocal = int(in.focal_loc.size() / 3), nrecv = int(in.recv.size() / 3), nsnap = int(in.snap.size() / 11);
    cgfd_ctx *c = nullptr;
    check(c, cgfd_create(&d, sizeof(Real) == 8 ? CGFD_FP64 : CGFD_FP32, 0, &c), "create");

    if (in.simplecar) check(c, cgfd_set_simplecar(c, 1), "simplecar");   // the SimpleCar build
    const void *drv[10];
    for (int i = 0; i < 10; ++i) drv[i] = in.metrics.data() + i * N;
    check(c, cgfd_set_metrics(c, drv), "metrics");                  // ParH2D derivF
    check(c, cgfd_set_media(c, in.media.data(), in.media.data() + N, in.media.data() + 2 * N), "media");
    const size_t len[3] = {NX, NY, NZ};
    const void *tabs[9];
    size_t off = 0;
    for (int t = 0; t < 9; ++t) {
        tabs[t] = in.pml.data() + off;
        off += len[t % 3];
    }
    check(c, cgfd_set_pml(c, in.nabs.data(), tabs), "pml");         // ParH2D apara
    if (nforce)                                                      // forceF
        check(c, cgfd_set_forces(c, nforce, in.force_loc.data(), in.force_vec.data(), in.force_stf.data(), nstf),
              "forces");
    if (nmoment)                                                     // momentF
        check(c, cgfd_set_moments(c, nmoment, in.moment_loc.data(), in.moment_vec.data(), in.moment_stf.data(), nstf),
              "moments");
    if (nfocal)                                                      // source_type 1
        check(c, cgfd_set_focal(c, nfocal, in.focal_loc.data(), in.focal_series.data(), in.focal_nt, in.focal_dt),
              "focal");
    if (nrecv) check(c, cgfd_set_receivers(c, nrecv, in.recv.data()), "receivers");   // C2DPointPick
    if (nsnap) check(c, cgfd_set_snapshots(c, nsnap, in.snap.data()), "snapshots");   // C2DSnapPick
    check(c, cgfd_set_pgv(c, in.pgv), "pgv");
    if (in.conindex > 0) check(c, cgfd_set_hybrid(c, in.conindex), "hybrid");
    check(c, cgfd_init(c), "init");                                 // VelCoeff

    check(c, cgfd_run(c, steps), "run");                            // the time loop
    check(c, cgfd_sync(c), "sync");

    std::vector<Real> traces(size_t(9) * nrecv * d.nt);             // PWgather
    if (nrecv) check(c, cgfd_get_seismo(c, traces.data()), "get_seismo");
    write_bin(out, "traces.bin", traces);
    std::vector<Real> pv(3 * NX * NY);                              // SynPV
    check(c, cgfd_get_pgv(c, pv.data()), "get_pgv");
    write_bin(out, "pgv.bin", pv);
    if (nrecv) {
        const bool meta = in.pnum.size() == size_t(nrecv);
        check(c, cgfd_write_seismo(c, (out + "/seismo.nc").c_str(), in.seismo_tinv, meta ? in.pnum.data() : nullptr,
                                   meta ? in.lnum.data() : nullptr, meta ? in.posx.data() : nullptr,
                                   meta ? in.posy.data() : nullptr, meta ? in.posz.data() : nullptr),
              "seismo.nc");
    }
    for (int i = 0; i < nsnap; ++i) {
        char name[32];
        std::snprintf(name, sizeof name, "/snap_%03d.nc", i + 1);
        check(c, cgfd_write_snap(c, i, (out + name).c_str()), "snap_NNN.nc");
    }
    if (in.pgv) check(c, cgfd_write_peakvel(c, (out + "/peakvel.nc").c_str(), 1), "peakvel.nc");
    cgfd_destroy(c);
    return 0;
}

template <typename Real>
int run_dir(const std::string &dir, const cgfd_dims &d, int prec, const int cnt[5], int pgv, const double extra[5],
            int steps) {
    Inputs<Real> in;
    in.d = d;
    in.pgv = pgv;
    const int nforce = cnt[0], nmoment = cnt[1], nrecv = cnt[2], nsnap = cnt[3], nfocal = cnt[4];
    in.focal_nt = int(extra[1]);
    in.focal_dt = extra[2];
    in.conindex = int(extra[3]);
    in.seismo_tinv = extra[4] > 0 ? int(extra[4]) : 1;
    const size_t NX = d.ni + 6, NY = d.nj + 6, NZ = d.nk + 6, N = NX * NY * NZ;
    const int nstf = 2 * (d.nt + 1);
    (void)prec;
    in.metrics = read_bin<Real>(dir, "metrics.bin", 10 * N);
    in.media = read_bin<Real>(dir, "media.bin", 3 * N);
    in.nabs = read_bin<int>(dir, "nabs.bin", 6);
    in.pml = read_bin<Real>(dir, "pml.bin", 3 * (NX + NY + NZ));
    in.force_loc = read_bin<int>(dir, "force_loc.bin", 3 * nforce);
    in.force_vec = read_bin<Real>(dir, "force_vec.bin", 3 * nforce);
    in.force_stf = read_bin<Real>(dir, "force_stf.bin", size_t(nforce) * nstf);
    in.moment_loc = read_bin<int>(dir, "moment_loc.bin", 3 * nmoment);
    in.moment_vec = read_bin<Real>(dir, "moment_vec.bin", 6 * nmoment);
    in.moment_stf = read_bin<Real>(dir, "moment_stf.bin", size_t(nmoment) * nstf);
    in.focal_loc = read_bin<int>(dir, "focal_loc.bin", 3 * nfocal);
    in.focal_series = read_bin<Real>(dir, "focal_series.bin", size_t(6) * nfocal * in.focal_nt);
    in.recv = read_bin<int>(dir, "recv.bin", 3 * nrecv);
    in.snap = read_bin<int>(dir, "snap.bin", 11 * nsnap);
    return run<Real>(in, steps, dir);
}

Inputs<float> from_conf(const cgfd_conf::Scenario &s) {
    Inputs<float> in;
    in.d = cgfd_dims{s.ni, s.nj, s.nk, (double)s.dh, (double)s.dt, s.nt};
    in.pgv = s.pgv;
    in.conindex = s.hybrid ? s.conindex : 0;
    for (int q = 0; q < 10; ++q) in.metrics.insert(in.metrics.end(), s.metrics[q].begin(), s.metrics[q].end());
    for (auto *a : {&s.alpha, &s.beta, &s.rho}) in.media.insert(in.media.end(), a->begin(), a->end());
    for (int q = 0; q < 9; ++q) in.pml.insert(in.pml.end(), s.pml[q].begin(), s.pml[q].end());
    in.nabs.assign(s.nabs, s.nabs + 6);
    in.force_loc = s.force_loc;
    in.force_vec = s.force_vec;
    in.force_stf = s.force_stf;
    in.moment_loc = s.moment_loc;
    in.moment_vec = s.moment_vec;
    in.moment_stf = s.moment_stf;
    in.focal_loc = s.focal_loc;
    in.focal_series = s.focal_series;
    in.focal_nt = s.focal_nt;
    in.focal_dt = s.focal_dt;
    in.recv = s.recv_loc;
    in.snap = s.snapshots;
    in.seismo_tinv = s.seismo_tinv;
    in.pnum = s.recv_pnum;
    in.lnum = s.recv_lnum;
    for (int i = 0; i < s.nrecv(); ++i) {
        in.posx.push_back(s.recv_pos[3 * i]);
        in.posy.push_back(s.recv_pos[3 * i + 1]);
        in.posz.push_back(s.recv_pos[3 * i + 2]);
    }
    return in;
}

void dump(const cgfd_conf::Scenario &s, const std::string &dir) {
    const Inputs<float> in = from_conf(s);
    FILE *f = std::fopen((dir + "/dims.txt").c_str(), "w");
    if (!f) die("cannot write " + dir + "/dims.txt");
    std::fprintf(f, "%d %d %d %.17g %.17g %d %d %d %d %d %d %d\n%d %d %.17g %d %d\n", s.ni, s.nj, s.nk, (double)s.dh,
                 (double)s.dt, s.nt, CGFD_FP32, s.nforce(), s.nmoment(), s.nrecv(), s.nsnap(), s.pgv, s.nfocal(),
                 s.focal_nt, (double)s.focal_dt, in.conindex, s.seismo_tinv);
    std::fclose(f);
    write_bin(dir, "metrics.bin", in.metrics);
    write_bin(dir, "media.bin", in.media);
    write_bin(dir, "nabs.bin", in.nabs);
    write_bin(dir, "pml.bin", in.pml);
    write_bin(dir, "force_loc.bin", in.force_loc);
    write_bin(dir, "force_vec.bin", in.force_vec);
    write_bin(dir, "force_stf.bin", in.force_stf);
    write_bin(dir, "moment_loc.bin", in.moment_loc);
    write_bin(dir, "moment_vec.bin", in.moment_vec);
    write_bin(dir, "moment_stf.bin", in.moment_stf);
    write_bin(dir, "focal_loc.bin", in.focal_loc);
    write_bin(dir, "focal_series.bin", in.focal_series);
    write_bin(dir, "recv.bin", in.recv);
    write_bin(dir, "snap.bin", in.snap);
    std::vector<float> crd(s.x);
    crd.insert(crd.end(), s.y.begin(), s.y.end());
    crd.insert(crd.end(), s.z.begin(), s.z.end());
    write_bin(dir, "coords.bin", crd);
    write_bin(dir, "recv_pos.bin", s.recv_pos);
    write_bin(dir, "recv_pnum.bin", s.recv_pnum);
    write_bin(dir, "recv_lnum.bin", s.recv_lnum);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2)
        die("usage: cgfd_child <dir> [steps] | cgfd_child --conf <SeisFD3D.conf> [--seispath <dir>] "
            "[--simplecar] (--dump <dir> | --out <dir> [--steps N])");
    if (std::strcmp(argv[1], "--conf") == 0) {
        std::string conf, seispath, dumpdir, outdir;
        int steps = -1, simplecar = 0;
        for (int a = 1; a < argc; ++a) {
            const std::string k = argv[a];
            if (k == "--simplecar") {
                simplecar = 1;
                continue;
            }
            if (a + 1 >= argc) die("missing value after " + k);
            const std::string v = argv[++a];
            if (k == "--conf") conf = v;
            else if (k == "--seispath") seispath = v;
            else if (k == "--dump") dumpdir = v;
            else if (k == "--out") outdir = v;
            else if (k == "--steps") steps = std::atoi(v.c_str());
            else die("unknown option " + k);
        }
        if (dumpdir.empty() == outdir.empty()) die("give exactly one of --dump and --out");
        cgfd_conf::Scenario s;
        try {
            s = cgfd_conf::load(conf, seispath);
        } catch (const cgfd_conf::ConfError &e) {
            die(e.what());
        }
        if (!dumpdir.empty()) {
            dump(s, dumpdir);
            return 0;
        }
        Inputs<float> in = from_conf(s);
        in.simplecar = simplecar;
        return run<float>(in, steps < 0 ? s.nt : steps, outdir);
    }
    const std::string dir = argv[1];
    FILE *f = std::fopen((dir + "/dims.txt").c_str(), "r");
    if (!f) die("cannot open " + dir + "/dims.txt");
    cgfd_dims d{};
    int prec = 0, cnt[5] = {0, 0, 0, 0, 0}, pgv = 1;
    int n = std::fscanf(f, "%d %d %d %lf %lf %d %d %d %d %d %d %d", &d.ni, &d.nj, &d.nk, &d.dh, &d.dt, &d.nt, &prec,
                        &cnt[0], &cnt[1], &cnt[2], &cnt[3], &pgv);
    double extra[5] = {0, 0, 0, 0, 0};
    int nfocal = 0;
    if (n == 12 && std::fscanf(f, "%d %lf %lf %lf %lf", &nfocal, &extra[1], &extra[2], &extra[3], &extra[4]) != 5)
        nfocal = 0, extra[1] = extra[2] = extra[3] = extra[4] = 0;
    cnt[4] = nfocal;
    std::fclose(f);
    if (n != 12) die("dims.txt: expected 12 values");
    const int steps = argc > 2 ? std::atoi(argv[2]) : d.nt;
    return prec == CGFD_FP64 ? run_dir<double>(dir, d, prec, cnt, pgv, extra, steps)
                             : run_dir<float>(dir, d, prec, cnt, pgv, extra, steps);
}
