// Drop-in for the reference's production GPU driver (hydrocodes/sph/gpu_ic.cpp):
//
//   dsph_gpu_ic <ic.bin> <dx> <t_end> [walltime] [--frac] [--g g] [--snap dt] [--damp f]
//               [--fp64 | --fast] [--materials basalt,iron,basalt] [--graph-steps k]
//               [--checkpoint file] [--checkpoint-every dt] [--restart file] [--max-steps n]
//
// Same arguments, IC format (gpu_ic.cpp:44-51), snapshot files gpu_snap_NNN.bin
// (gpu_ic.cpp:118-123; written here in IC order), progress lines with the
// "melt=X%" token that boundary_daemon.py:240-248 parses, and the closing DONE /
// PROFILE lines.  Physics constants follow gpu_ic.cpp:54-59,67: eta = 1.3,
// h_init = 1.3 dx, alpha = 1, beta = 2, eps = 0.01, cfl = 0.2, and with --frac
// artificial stress 0.3 plus Weibull damage seeded from mt19937(12345).  The
// step itself follows the CPU solver (sph.hpp:357-377): two density iterations
// per round, not the single iteration of the Metal port.
//
// Checkpoints use run_ic's layout (run_ic.cpp:36-55: header [xlo xhi ylo yhi zlo
// zhi n t dx], rows [x y z vx vy vz mass u mat fixed D eps_act S(xx xy xz yy yz
// zz)], written to <file>.tmp and renamed) extended by rows [h ax ay az dudt
// dSdt(6)] and a trailer [20241204 steps dt_next], which makes a restart continue
// bit for bit.  --restart accepts both forms; a plain run_ic checkpoint restores
// what it holds and re-runs init(), as run_ic.cpp:76-102 does.  With --restart the
// positional IC file is not read.
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

constexpr double kExtMagic = 20241204.0;
constexpr double kExtMagicV0 = 20241203.0;
// file tensor order {xx,xy,xz,yy,yz,zz} <-> memory order {xx,yy,zz,xy,xz,yz}
constexpr int kFileToMem[6] = {0, 3, 5, 1, 2, 4};
constexpr int kMemToFile[6] = {0, 3, 4, 1, 5, 2};

void write_checkpoint(dsph::System& S, const std::string& path, const double lo[3], const double hi[3], double dx) {
    const dsph_diag d = S.diagnostics();
    std::vector<double> pos = S.get(DSPH_F_POS, 3), vel = S.get(DSPH_F_VEL, 3), mass = S.get(DSPH_F_MASS),
                        u = S.get(DSPH_F_U), D = S.get(DSPH_F_D), ea = S.get(DSPH_F_EPS_ACT),
                        Sm = S.get(DSPH_F_S, 6), h = S.get(DSPH_F_HH), acc = S.get(DSPH_F_ACC, 3),
                        dudt = S.get(DSPH_F_DUDT), dS = S.get(DSPH_F_DSDT, 6);
    std::vector<int32_t> mat = S.get_mat();
    std::vector<uint8_t> fixed = S.get_fixed();
    const size_t n = mass.size();
    std::vector<double> buf;
    buf.reserve(9 + 29 * n + 3);
    for (int a = 0; a < 3; a++) { buf.push_back(lo[a]); buf.push_back(hi[a]); }
    buf.push_back((double)n); buf.push_back(d.t); buf.push_back(dx);
    for (size_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) buf.push_back(pos[3 * i + a]);
        for (int a = 0; a < 3; a++) buf.push_back(vel[3 * i + a]);
        buf.push_back(mass[i]); buf.push_back(u[i]); buf.push_back((double)mat[i]); buf.push_back(fixed[i] ? 1.0 : 0.0);
        buf.push_back(D[i]); buf.push_back(ea[i]);
        for (int k = 0; k < 6; k++) buf.push_back(Sm[6 * i + kMemToFile[k]]);
    }
    for (size_t i = 0; i < n; i++) {
        buf.push_back(h[i]);
        for (int a = 0; a < 3; a++) buf.push_back(acc[3 * i + a]);
        buf.push_back(dudt[i]);
        for (int k = 0; k < 6; k++) buf.push_back(dS[6 * i + kMemToFile[k]]);
    }
    buf.push_back(kExtMagic); buf.push_back((double)d.steps); buf.push_back(d.dt_next);
    const std::string tmp = path + ".tmp";
    FILE* o = std::fopen(tmp.c_str(), "wb");
    if (!o || std::fwrite(buf.data(), 8, buf.size(), o) != buf.size()) {
        std::fprintf(stderr, "cannot write %s\n", tmp.c_str());
        std::exit(1);
    }
    std::fclose(o);
    if (std::rename(tmp.c_str(), path.c_str()) != 0) { std::fprintf(stderr, "cannot rename %s\n", tmp.c_str()); std::exit(1); }
}

struct Checkpoint {
    double hdr[9];
    long n = 0;
    std::vector<double> rows, ext;   // n x 18, n x 11 (ext empty for a plain checkpoint)
    double steps = 0, dt_next = 0;
};

bool read_checkpoint(const std::string& path, Checkpoint& ck) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<double> raw((size_t)bytes / 8);
    const bool ok = std::fread(raw.data(), 8, raw.size(), f) == raw.size();
    std::fclose(f);
    if (!ok || raw.size() < 9) return false;
    std::copy(raw.begin(), raw.begin() + 9, ck.hdr);
    ck.n = std::llround(ck.hdr[6]);
    const size_t body = 9 + 18 * (size_t)ck.n, rest = raw.size() - std::min(raw.size(), body);
    if (raw.size() < body) return false;
    ck.rows.assign(raw.begin() + 9, raw.begin() + body);
    if (rest == 11 * (size_t)ck.n + 3 && raw[raw.size() - 3] == kExtMagic) {
        ck.steps = raw[raw.size() - 2];
        ck.dt_next = raw[raw.size() - 1];
    } else if (rest == 11 * (size_t)ck.n + 2 && raw[raw.size() - 2] == kExtMagicV0) {
        ck.steps = raw[raw.size() - 1];
    } else if (rest != 0) {
        return false;
    }
    if (rest) ck.ext.assign(raw.begin() + body, raw.begin() + body + 11 * (size_t)ck.n);
    return true;
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
    bool frac = false, fp64 = false, fast = false;
    double gz = 0.0, snap_dt = 0.05, damp = 1.0;
    int graph_steps = 25;
    std::string mats = "basalt,iron,basalt";       // gpu_ic.cpp:67 (material 2 = trackable basalt impactor)
    std::string ckpt_path, restart_path;
    double ckpt_every = 0.0;
    long max_steps = -1;
    std::vector<char*> av;
    for (int i = 0; i < argc; i++) {
        std::string s = argv[i];
        if (s == "--frac") frac = true;
        else if (s == "--fp64") fp64 = true;
        else if (s == "--fast") fast = true;     // DSPH_MIXED_FAST: FP32 density sums too
        else if (s == "--g" && i + 1 < argc) gz = std::atof(argv[++i]);
        else if (s == "--snap" && i + 1 < argc) snap_dt = std::atof(argv[++i]);
        else if (s == "--damp" && i + 1 < argc) damp = std::atof(argv[++i]);
        else if (s == "--materials" && i + 1 < argc) mats = argv[++i];
        else if (s == "--graph-steps" && i + 1 < argc) graph_steps = std::atoi(argv[++i]);
        else if (s == "--checkpoint" && i + 1 < argc) ckpt_path = argv[++i];
        else if (s == "--checkpoint-every" && i + 1 < argc) ckpt_every = std::atof(argv[++i]);
        else if (s == "--restart" && i + 1 < argc) restart_path = argv[++i];
        else if (s == "--max-steps" && i + 1 < argc) max_steps = std::atol(argv[++i]);
        else av.push_back(argv[i]);
    }
    if (av.size() < 4) {
        std::printf("usage: dsph_gpu_ic ic.bin dx t_end [walltime] [--frac] [--g g_mars] [--snap dt] [--damp f]"
                    " [--fp64 | --fast] [--materials a,b,c] [--checkpoint file] [--checkpoint-every dt] [--restart file]"
                    " [--max-steps n]\n");
        return 1;
    }
    const char* icf = av[1];
    const double dx = std::atof(av[2]), t_end = std::atof(av[3]);
    const double walltime = av.size() > 4 ? std::atof(av[4]) : 1e30;

    dsph::System S;
    S.materials = parse_materials(mats);
    S.eta = 1.3; S.h_init = 1.3 * dx; S.alpha = 1.0; S.beta = 2.0; S.eps = 0.01; S.cfl = 0.2;
    S.eps_as = frac ? 0.3 : 0.0; S.damage_on = frac;
    S.g = gz; S.damp = damp; S.t_end = t_end;
    S.precision = fp64 ? DSPH_FP64 : (fast ? DSPH_MIXED_FAST : DSPH_MIXED);
    auto add_row = [&](long i, const double* r) -> bool {
        S.cur_mat = (int)r[8];
        if (S.cur_mat < 0 || S.cur_mat >= (int)S.materials.size()) {
            std::printf("particle %ld: material %d not in the material list\n", i, S.cur_mat);
            return false;
        }
        S.add({r[0], r[1], r[2]}, {r[3], r[4], r[5]}, r[6], r[7], r[9] != 0.0);
        return true;
    };
    double dom_lo[3], dom_hi[3];
    long n = 0;
    int n_proj = 0;
    double t0 = 0.0;
    Checkpoint ck;
    const bool restart = !restart_path.empty();
    if (restart) {
        // ---- restart from a checkpoint (run_ic layout, optionally extended) ----
        if (!read_checkpoint(restart_path, ck)) { std::printf("cannot read checkpoint %s\n", restart_path.c_str()); return 1; }
        n = ck.n;
        t0 = ck.hdr[7];
        for (int a = 0; a < 3; a++) { dom_lo[a] = ck.hdr[2 * a]; dom_hi[a] = ck.hdr[2 * a + 1]; }
        S.set_domain(dom_lo[0], dom_hi[0], false, dom_lo[1], dom_hi[1], false, dom_lo[2], dom_hi[2], false);
        for (long i = 0; i < n; i++) {
            const double* r = &ck.rows[18 * (size_t)i];
            if (!add_row(i, r)) return 1;
            if (r[8] >= 1) n_proj++;
        }
    } else {
        // ---- load IC (run_ic format) ----
        FILE* f = std::fopen(icf, "rb");
        if (!f) { std::printf("cannot open %s\n", icf); return 1; }
        double hdr[7];
        if (std::fread(hdr, 8, 7, f) != 7) { std::printf("bad header in %s\n", icf); return 1; }
        n = std::llround(hdr[6]);
        std::vector<double> rows(10 * (size_t)n);
        if ((long)std::fread(rows.data(), 80, n, f) != n) { std::printf("short read in %s\n", icf); return 1; }
        std::fclose(f);
        for (int a = 0; a < 3; a++) { dom_lo[a] = hdr[2 * a]; dom_hi[a] = hdr[2 * a + 1]; }
        S.set_domain(dom_lo[0], dom_hi[0], false, dom_lo[1], dom_hi[1], false, dom_lo[2], dom_hi[2], false);
        for (long i = 0; i < n; i++) {
            const double* r = &rows[10 * (size_t)i];
            if (!add_row(i, r)) return 1;
            if (r[8] >= 1) n_proj++;
        }
    }
    std::printf("loaded %s: n=%ld (%d projectile)  dx=%.0f  frac=%d  g=%.2f  precision=%s\n",
                restart ? restart_path.c_str() : icf, n, n_proj, dx, frac, gz,
                fp64 ? "fp64" : (fast ? "mixed_fast" : "mixed"));
    try {
        if (restart) {
            std::vector<double> D(n), ea(n), Sm(6 * (size_t)n);
            for (long i = 0; i < n; i++) {
                const double* r = &ck.rows[18 * (size_t)i];
                D[i] = r[10];
                ea[i] = r[11];
                for (int k = 0; k < 6; k++) Sm[6 * i + k] = r[12 + kFileToMem[k]];
            }
            S.set(DSPH_F_D, D);
            S.set(DSPH_F_EPS_ACT, ea);
            S.set(DSPH_F_S, Sm);
            if (!ck.ext.empty()) {
                std::vector<double> h(n), acc(3 * (size_t)n), dudt(n), dS(6 * (size_t)n);
                for (long i = 0; i < n; i++) {
                    const double* e = &ck.ext[11 * (size_t)i];
                    h[i] = e[0];
                    for (int a = 0; a < 3; a++) acc[3 * i + a] = e[1 + a];
                    dudt[i] = e[4];
                    for (int k = 0; k < 6; k++) dS[6 * i + k] = e[5 + kFileToMem[k]];
                }
                S.set(DSPH_F_HH, h);
                S.set(DSPH_F_ACC, acc);
                S.set(DSPH_F_DUDT, dudt);
                S.set(DSPH_F_DSDT, dS);
                S.set_clock(t0, (int64_t)std::llround(ck.steps), ck.dt_next);
            } else {
                S.set_clock(t0, 0, 0.0);
                S.init();                 // a plain run_ic checkpoint: re-derive h, rho and forces
            }
            std::printf("restart at t=%.6f step=%lld (%s checkpoint)\n", t0, (long long)std::llround(ck.steps),
                        ck.ext.empty() ? "plain" : "extended");
        } else {
            if (frac) S.seed_damage(dx, 12345);
            S.init();
        }
    } catch (const dsph::Error& e) {
        std::printf("init failed: %s\n", e.what());
        return 1;
    }
    const dsph_diag d0 = S.diagnostics();
    const double E0 = d0.e_tot;
    std::printf("init done, E0=%.4e\n", E0);
    std::fflush(stdout);

    Snapshotter snap;
    double t = d0.t;
    long step = d0.steps;       // global step count (continues across a restart)
    int isnap = (int)std::floor(t / snap_dt + 1e-9);
    double next_snap = (isnap + 1) * snap_dt;
    double next_ckpt = ckpt_every > 0 ? (std::floor(t / ckpt_every + 1e-9) + 1) * ckpt_every : 1e300;
    const long step0 = step;
    long last_report = step;
    auto checkpoint = [&]() {
        if (ckpt_path.empty()) return;
        write_checkpoint(S, ckpt_path, dom_lo, dom_hi, dx);
        std::printf("checkpoint %s at t=%.6f step=%ld\n", ckpt_path.c_str(), t, step);
        std::fflush(stdout);
    };
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
        std::printf("  t=%.4f step=%ld proj: x=%.0f z=%.0f |v|=%.0f vz=%.0f melt=%.0f%% umax=%.1e [%.0fs]\n", tt, step,
                    cx / ni, cz / ni, sp / ni, vz / ni, 100.0 * nm / ni, um, now() - wall0);
        std::fflush(stdout);
        T_report += now() - a;
    };
    try {
        while (t < t_end && (max_steps < 0 || step < max_steps)) {
            // device CFL step (sph.hpp:350-355), clamped to land on t_end (gpu_ic.cpp:130)
            const double dtc = S.timestep();     // raises the device's trouble flags
            if (!(dtc > 0.0) || !std::isfinite(dtc)) {
                std::printf("step %ld failed: device time step %g is not positive and finite\n", step, dtc);
                return 1;
            }
            const double a = now();
            const double t_prev = t;
            if (t + dtc > t_end) {
                S.step(t_end - t);
                t = t_end;
                step++;
            } else {
                // a few steps per host round trip; stop short of t_end and snapshot times
                int k = 1;
                const double horizon = std::min(std::min(t_end, next_snap), next_ckpt) - t;
                if (dtc > 0 && horizon > 4 * dtc) k = std::max(1, std::min(graph_steps, (int)(horizon / dtc) - 2));
                if (max_steps >= 0) k = (int)std::min<long>(k, max_steps - step);
                S.run(k, 0.0, true);             // raises the device's trouble flags
                dsph_diag d = S.diagnostics();
                t = d.t;
                step += k;
            }
            if (!(t > t_prev)) {
                std::printf("step %ld failed: the clock did not advance (t=%.9g)\n", step, t);
                return 1;
            }
            T_step += now() - a;
            if (step / 25 != last_report / 25 || t >= t_end) { report(t); last_report = step; }
            if (t >= next_snap) { snap.write(S, isnap++); next_snap += snap_dt; }
            if (t >= next_ckpt) { checkpoint(); next_ckpt += ckpt_every; }
            if (now() - wall0 > walltime) { std::printf("walltime hit\n"); break; }
        }
        checkpoint();
    } catch (const dsph::Error& e) {
        std::printf("step %ld failed: %s\n", step, e.what());
        return 1;
    }
    const dsph_diag d1 = S.diagnostics();
    const long taken = std::max(step - step0, 1L);
    std::printf("DONE: %ld steps to t=%.4f  %.1f ms/step  E drift=%.2e\n", step, t, 1000 * (now() - wall0) / taken,
                (d1.e_tot - E0) / E0);
    std::printf("PROFILE (ms/step): step=%.1f report=%.1f\n", 1000 * T_step / taken, 1000 * T_report / taken);
    return 0;
}
