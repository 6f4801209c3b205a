// Drop-in for the reference's production GPU driver (hydrocodes/sph/gpu_ic.cpp):
//
//   dsph_gpu_ic <ic.bin> <dx> <t_end> [walltime] [--frac] [--g g] [--snap dt] [--damp f]
//               [--fp64] [--materials basalt,iron,basalt] [--graph-steps k]
//
// Same arguments, IC format (gpu_ic.cpp:44-51), snapshot files gpu_snap_NNN.bin
// (gpu_ic.cpp:118-123; written here in IC order), progress lines with the
// "melt=X%" token that boundary_daemon.py:240-248 parses, and the closing DONE /
// PROFILE lines.  Physics constants follow gpu_ic.cpp:54-59,67: eta = 1.3,
// h_init = 1.3 dx, alpha = 1, beta = 2, eps = 0.01, cfl = 0.2, and with --frac
// artificial stress 0.3 plus Weibull damage seeded from mt19937(12345).  The
// step itself follows the CPU solver (sph.hpp:357-377): two density iterations
// per round, not the single iteration of the Metal port.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dsph_system.hpp"

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

dsph_material material_by_name(const std::string& s) {
    if (s == "iron") return dsph::iron();
    if (s == "basalt") return dsph::basalt();
    if (s == "water") return dsph::water();
    if (s == "granite")   // the impact workloads' granite (Tillotson with Weibull k=1e39, m=9)
        return dsph::tillotson(2680, 18e9, 18e9, 0.5, 1.3, 5, 5, 16e6, 3.5e6, 18e6, 30e9, 2e9, 9.0, 1e39);
    if (s == "iron80") return dsph::tillotson(7800, 128e9, 105e9, 0.5, 1.5, 5, 5, 9.5e6, 2.4e6, 8.67e6, 80e9, 1e9);
    std::fprintf(stderr, "unknown material %s\n", s.c_str());
    std::exit(1);
}

std::vector<dsph_material> parse_materials(const std::string& list) {
    std::vector<dsph_material> out;
    size_t b = 0;
    while (b <= list.size()) {
        size_t e = list.find(',', b);
        if (e == std::string::npos) e = list.size();
        out.push_back(material_by_name(list.substr(b, e - b)));
        b = e + 1;
    }
    return out;
}

struct Snapshotter {
    void write(dsph::System& S, int idx) {
        char fn[64];
        std::snprintf(fn, sizeof(fn), "gpu_snap_%03d.bin", idx);
        std::vector<double> pos = S.get(DSPH_F_POS, 3), vel = S.get(DSPH_F_VEL, 3), u = S.get(DSPH_F_U),
                            rho = S.get(DSPH_F_RHO), h = S.get(DSPH_F_HH);
        std::vector<int32_t> mat = S.get_mat();
        const long n = (long)u.size();
        std::vector<double> rec(10 * (size_t)n);
        for (long i = 0; i < n; i++) {
            double* r = &rec[10 * (size_t)i];
            r[0] = pos[3 * i]; r[1] = pos[3 * i + 1]; r[2] = pos[3 * i + 2];
            r[3] = vel[3 * i]; r[4] = vel[3 * i + 1]; r[5] = vel[3 * i + 2];
            r[6] = u[i]; r[7] = rho[i]; r[8] = h[i]; r[9] = (double)mat[i];
        }
        FILE* o = std::fopen(fn, "wb");
        if (!o) { std::fprintf(stderr, "cannot write %s\n", fn); std::exit(1); }
        int64_t N = n;
        std::fwrite(&N, 8, 1, o);
        std::fwrite(rec.data(), 8, rec.size(), o);
        std::fclose(o);
    }
};

}  // namespace

int main(int argc, char** argv) {
    bool frac = false, fp64 = false;
    double gz = 0.0, snap_dt = 0.05, damp = 1.0;
    int graph_steps = 25;
    std::string mats = "basalt,iron,basalt";       // gpu_ic.cpp:67 (material 2 = trackable basalt impactor)
    std::vector<char*> av;
    for (int i = 0; i < argc; i++) {
        std::string s = argv[i];
        if (s == "--frac") frac = true;
        else if (s == "--fp64") fp64 = true;
        else if (s == "--g" && i + 1 < argc) gz = std::atof(argv[++i]);
        else if (s == "--snap" && i + 1 < argc) snap_dt = std::atof(argv[++i]);
        else if (s == "--damp" && i + 1 < argc) damp = std::atof(argv[++i]);
        else if (s == "--materials" && i + 1 < argc) mats = argv[++i];
        else if (s == "--graph-steps" && i + 1 < argc) graph_steps = std::atoi(argv[++i]);
        else av.push_back(argv[i]);
    }
    if (av.size() < 4) {
        std::printf("usage: dsph_gpu_ic ic.bin dx t_end [walltime] [--frac] [--g g_mars] [--snap dt] [--damp f]"
                    " [--fp64] [--materials a,b,c]\n");
        return 1;
    }
    const char* icf = av[1];
    const double dx = std::atof(av[2]), t_end = std::atof(av[3]);
    const double walltime = av.size() > 4 ? std::atof(av[4]) : 1e30;

    // ---- load IC (run_ic format) ----
    FILE* f = std::fopen(icf, "rb");
    if (!f) { std::printf("cannot open %s\n", icf); return 1; }
    double hdr[7];
    if (std::fread(hdr, 8, 7, f) != 7) { std::printf("bad header in %s\n", icf); return 1; }
    const long n = std::llround(hdr[6]);
    std::vector<double> rows(10 * (size_t)n);
    if ((long)std::fread(rows.data(), 80, n, f) != n) { std::printf("short read in %s\n", icf); return 1; }
    std::fclose(f);

    dsph::System S;
    S.materials = parse_materials(mats);
    S.eta = 1.3; S.h_init = 1.3 * dx; S.alpha = 1.0; S.beta = 2.0; S.eps = 0.01; S.cfl = 0.2;
    S.eps_as = frac ? 0.3 : 0.0; S.damage_on = frac;
    S.g = gz; S.damp = damp; S.t_end = t_end;
    S.precision = fp64 ? DSPH_FP64 : DSPH_MIXED;
    S.set_domain(hdr[0], hdr[1], false, hdr[2], hdr[3], false, hdr[4], hdr[5], false);
    int n_proj = 0;
    for (long i = 0; i < n; i++) {
        const double* r = &rows[10 * (size_t)i];
        S.cur_mat = (int)r[8];
        if (S.cur_mat < 0 || S.cur_mat >= (int)S.materials.size()) {
            std::printf("particle %ld: material %d not in the material list\n", i, S.cur_mat);
            return 1;
        }
        S.add({r[0], r[1], r[2]}, {r[3], r[4], r[5]}, r[6], r[7], r[9] != 0.0);
        if (S.cur_mat >= 1) n_proj++;
    }
    std::printf("loaded %s: n=%ld (%d projectile)  dx=%.0f  frac=%d  g=%.2f  precision=%s\n", icf, n, n_proj, dx,
                frac, gz, fp64 ? "fp64" : "mixed");
    try {
        if (frac) S.seed_damage(dx, 12345);
        S.init();
    } catch (const dsph::Error& e) {
        std::printf("init failed: %s\n", e.what());
        return 1;
    }
    const dsph_diag d0 = S.diagnostics();
    const double E0 = d0.e_tot;
    std::printf("init done, E0=%.4e\n", E0);
    std::fflush(stdout);

    Snapshotter snap;
    double t = 0.0, next_snap = snap_dt;
    int step = 0, isnap = 0, last_report = 0;
    const double wall0 = now();
    double T_step = 0, T_report = 0;
    std::vector<int32_t> mat = S.get_mat();
    auto report = [&](double tt) {
        const double a = now();
        std::vector<double> P = S.get(DSPH_F_POS, 3), V = S.get(DSPH_F_VEL, 3), U = S.get(DSPH_F_U);
        double cx = 0, cz = 0, vz = 0, sp = 0, um = 0;
        int ni = 0, nm = 0;
        for (long i = 0; i < n; i++) {
            if (mat[i] < 1) continue;        // projectile: material >= 1 (gpu_ic.cpp:114)
            cx += P[3 * i]; cz += P[3 * i + 2]; vz += V[3 * i + 2];
            sp += std::sqrt(V[3 * i] * V[3 * i] + V[3 * i + 1] * V[3 * i + 1] + V[3 * i + 2] * V[3 * i + 2]);
            if (U[i] > 1e6) nm++;
            um = std::max(um, U[i]);
            ni++;
        }
        ni = ni > 0 ? ni : 1;
        std::printf("  t=%.4f step=%d proj: x=%.0f z=%.0f |v|=%.0f vz=%.0f melt=%.0f%% umax=%.1e [%.0fs]\n", tt, step,
                    cx / ni, cz / ni, sp / ni, vz / ni, 100.0 * nm / ni, um, now() - wall0);
        std::fflush(stdout);
        T_report += now() - a;
    };
    try {
        while (t < t_end) {
            // device CFL step (sph.hpp:350-355), clamped to land on t_end (gpu_ic.cpp:130)
            const double dtc = S.timestep();
            const double a = now();
            if (t + dtc > t_end) {
                S.step(t_end - t);
                t = t_end;
                step++;
            } else {
                // a few steps per host round trip; stop short of t_end and snapshot times
                int k = 1;
                const double horizon = std::min(t_end, next_snap) - t;
                if (dtc > 0 && horizon > 4 * dtc) k = std::max(1, std::min(graph_steps, (int)(horizon / dtc) - 2));
                S.run(k, 0.0, true);
                dsph_diag d = S.diagnostics();
                t = d.t;
                step += k;
            }
            T_step += now() - a;
            if (step / 25 != last_report / 25 || t >= t_end) { report(t); last_report = step; }
            if (t >= next_snap) { snap.write(S, isnap++); next_snap += snap_dt; }
            if (now() - wall0 > walltime) { std::printf("walltime hit\n"); break; }
        }
    } catch (const dsph::Error& e) {
        std::printf("step %d failed: %s\n", step, e.what());
        return 1;
    }
    const dsph_diag d1 = S.diagnostics();
    std::printf("DONE: %d steps to t=%.4f  %.1f ms/step  E drift=%.2e\n", step, t, 1000 * (now() - wall0) / std::max(step, 1),
                (d1.e_tot - E0) / E0);
    std::printf("PROFILE (ms/step): step=%.1f report=%.1f\n", 1000 * T_step / std::max(step, 1),
                1000 * T_report / std::max(step, 1));
    return 0;
}
