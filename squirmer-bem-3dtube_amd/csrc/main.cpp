// main.cpp -- squirmer_bem, the command-line front end of libsqbem (one MI355X).
//
//   squirmer_bem <input-file> [key=value ...]
//
// The input file (and the key=value overrides after it, applied in order) selects a mode
// and its parameters. Modes, keys, defaults, output files and their CSV schemas are those
// of the reference driver (src/main.cpp:27-201); the console messages are this driver's.
//
//   duct        pressure-driven tube flow            flow_rate.csv, panels.csv, profile.csv
//   kinematics  U, Omega per (a/R, beta), towed drag kinematics.csv
//   field       velocity in the plane y = 0          field_B1.csv, field_B2.csv, field_info.csv
//   trajectory  time-integrated path                 <name>.csv (<name>.ckpt while paused)
//
// Keys beyond the reference's: precision = fp64 | fp32 | fp16 | bf16 (factor precision;
// the refined modes return FP64-accurate results by iterative refinement) and n_r (rings
// per tube cap of the swimmer modes; the reference's Tube default is 4).
//
// Kinematics sweeps go through the batched solver: every configuration without a towed
// problem is solved in one sqbem_batch_solve call (tubes shared by content, both squirming
// modes as two instances of one configuration); the on-axis configurations whose towed drag
// is asked for are solved one at a time, since the towed problem needs their solver.
#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "duct_flow.hpp"
#include "swimmer.hpp"
#include "trajectory.hpp"
#include "utils/config.hpp"
#include "utils/io.hpp"

using namespace bem;

namespace {

int factor_precision(const Config& cfg) {
    const std::string name = cfg.str("precision", "fp64");
    const std::pair<const char*, int> known[] = {
        {"fp64", SQBEM_FP64}, {"fp32", SQBEM_FP32_IR}, {"fp16", SQBEM_FP16_IR}, {"bf16", SQBEM_BF16_IR}};
    for (const auto& k : known)
        if (name == k.first) return k.second;
    throw std::runtime_error("unknown precision '" + name + "' (fp64|fp32|fp16|bf16)");
}

// ---------------------------------------------------------------- duct
void duct_mode(const Config& cfg, const std::string& out) {
    const double R = cfg.num("R", 1.0), L = cfg.num("L", 4.0), mu = cfg.num("mu", 1.0);
    const double p_out = cfg.num("p_out", 0.0), dP_detail = cfg.num("dP_detail", 1.0);
    const int n_profile = cfg.integer("n_profile", 20);
    Timer clock;
    const DuctFlow duct(R, L, cfg.integer("n_theta", 24), cfg.integer("n_z", 24), cfg.integer("n_r", 6), mu);
    std::printf("[duct] %d panels (%d unknowns) assembled and factored: %.2f s\n", duct.mesh.P(), 3 * duct.mesh.P(),
                clock.seconds());

    Csv rates(out + "/flow_rate.csv", "dP,Q,Q_exact,rel_err");
    for (const double dP : cfg.list("dP_list", {0.5, 1, 2, 4})) {
        const double Q = duct.solve(p_out + dP, p_out).Q, exact = duct.exact_Q(dP);
        const double err = std::fabs(Q - exact) / exact;
        rates.row({dP, Q, exact, err});
        std::printf("[duct] dP %g: Q %.6f (Poiseuille %.6f, error %.1e)\n", dP, Q, exact, err);
    }

    const DuctFlow::Solution detail = duct.solve(p_out + dP_detail, p_out);
    write_panels(out + "/panels.csv", duct.mesh, detail.u, detail.f);
    // radial lines at three azimuths through the mid-plane z = L/2, out to 0.95 R
    std::vector<V3> pts;
    for (const double phi : {0.0, PI / 5, PI / 2})
        for (int k = 0; k < n_profile; ++k) {
            const double r = 0.95 * R * k / (n_profile - 1);
            pts.push_back(V3{r * std::cos(phi), r * std::sin(phi), L / 2});
        }
    const std::vector<V3> vel = duct.velocity(detail, pts);
    Csv profile(out + "/profile.csv", "x,y,z,r,ux,uy,uz,uz_exact",
                {"interior velocity at z = L/2, dP = " + std::to_string(dP_detail)});
    for (size_t k = 0; k < pts.size(); ++k) {
        const double r = std::hypot(pts[k].x, pts[k].y);
        profile.row({pts[k].x, pts[k].y, pts[k].z, r, vel[k].x, vel[k].y, vel[k].z, duct.exact_uz(r, dP_detail)});
    }
    std::printf("[duct] %s/flow_rate.csv, panels.csv, profile.csv: %.2f s\n", out.c_str(), clock.seconds());
}

// ---------------------------------------------------------------- kinematics
struct KinematicsRow {
    double a_over_R, beta;
    int panels = 0;
    V3 U[2], Om[2];  // squirming modes B1 = 1 and B2 = 1
    double drag_K = NAN, seconds = 0;
};

void kinematics_mode(const Config& cfg, const std::string& out) {
    const double a = 1.0;  // sphere radius: lengths in units of a, R = a / (a/R)
    const std::vector<double> lams = cfg.list("a_over_R", {0.3}), betas = cfg.list("beta", {0.0});
    const int n_theta = cfg.integer("n_theta", 24), level = cfg.integer("sphere_level", 3);
    const int n_r = cfg.integer("n_r", 4), prec = factor_precision(cfg);
    const double L_over_R = cfg.num("L_over_R", 4.0);
    const bool want_drag = cfg.integer("towed_drag", 1) != 0;
    const V3 ez{0, 0, 1};

    std::vector<KinematicsRow> rows;
    // batched configurations: meshes kept alive for the call
    std::vector<std::unique_ptr<Mesh>> meshes;
    std::vector<BatchConfig> batch;
    std::vector<size_t> batch_row;  // row of each batched configuration (two entries per row)
    for (const double lam : lams) {
        const double R = lam > 0 ? a / lam : 0.0;
        std::unique_ptr<Tube> tube;        // for the rows solved one at a time
        const Mesh* tube_mesh = nullptr;   // for the batched rows
        for (const double beta : betas) {
            KinematicsRow row{lam, beta};
            const double b = lam > 0 ? beta * (R - a) : 0.0;
            const V3 c{b, 0, 0};
            const bool towed = want_drag && std::fabs(b) < 1e-12;
            if (lam > 0 && !towed) {
                if (!tube_mesh) {
                    meshes.push_back(std::make_unique<Mesh>(swimmer_tube_mesh(R, L_over_R * R, a, n_theta, n_r)));
                    tube_mesh = meshes.back().get();
                }
                meshes.push_back(std::make_unique<Mesh>(icosphere(a, level, c)));
                const Mesh* sphere = meshes.back().get();
                row.panels = tube_mesh->P() + sphere->P();
                for (int m = 0; m < 2; ++m) {
                    batch.push_back({tube_mesh, sphere, c, ez, m == 0 ? 1.0 : 0.0, m == 0 ? 0.0 : 1.0});
                    batch_row.push_back(rows.size());
                }
            } else {  // unbounded, or the towed problem is wanted: one solver for the row
                Timer clock;
                if (lam > 0 && !tube) tube = std::make_unique<Tube>(R, L_over_R * R, a, n_theta, n_r, 1.0, prec);
                SwimmerSolver s(tube.get(), icosphere(a, level, c), c, 1.0, prec);
                const auto m = s.solve_free({s.slip(ez, 1, 0), s.slip(ez, 0, 1)});
                for (int k = 0; k < 2; ++k) {
                    row.U[k] = m[k].U;
                    row.Om[k] = m[k].Om;
                }
                if (towed) row.drag_K = -s.force(s.solve_prescribed(ez, V3{})).z / (6 * PI);
                row.panels = s.Ns + s.Nt;
                row.seconds = clock.seconds();
            }
            rows.push_back(row);
        }
    }
    if (!batch.empty()) {
        Timer clock;
        const BatchResult r = batch_solve(batch, 1.0, prec);
        const double share = clock.seconds() / (double)(batch.size() / 2);  // per configuration
        for (size_t j = 0; j < batch.size(); ++j) {
            KinematicsRow& row = rows[batch_row[j]];
            const int m = batch[j].B1 != 0 ? 0 : 1;
            row.U[m] = r.U[j];
            row.Om[m] = r.Om[j];
            row.seconds = share;
        }
        std::printf("[kinematics] %zu configurations in one batched solve: %.2f s\n", batch.size() / 2,
                    clock.seconds());
    }

    Csv table(out + "/kinematics.csv",
              "a_over_R,beta,panels,UB1x,UB1y,UB1z,OB1x,OB1y,OB1z,UB2x,UB2y,UB2z,OB2x,OB2y,OB2z,drag_K,seconds",
              {"U, Omega for B1=1 (neutral mode) and B2=1 (dipole mode); a=1, mu=1, e=z. "
               "Any squirmer: U = U_B1 + alpha*U_B2. drag_K = axial drag/(6 pi mu a U), sphere on axis only. "
               "a_over_R=0 means unbounded fluid."});
    const double U0 = 2.0 / 3.0;  // free-space speed of the B1 = 1 squirmer
    for (const KinematicsRow& w : rows) {
        table.row({w.a_over_R, w.beta, (double)w.panels, w.U[0].x, w.U[0].y, w.U[0].z, w.Om[0].x, w.Om[0].y,
                   w.Om[0].z, w.U[1].x, w.U[1].y, w.U[1].z, w.Om[1].x, w.Om[1].y, w.Om[1].z, w.drag_K, w.seconds});
        std::printf("[kinematics] a/R %.2f beta %.2f: U_B1/U0 %.5f, Om_B1 (%+.2e, %+.2e, %+.2e), U_B2/U0 %+.5f, "
                    "K %.4f\n",
                    w.a_over_R, w.beta, w.U[0].z / U0, w.Om[0].x, w.Om[0].y, w.Om[0].z, w.U[1].z / U0, w.drag_K);
    }
}

// ---------------------------------------------------------------- field
void field_mode(const Config& cfg, const std::string& out) {
    const double a = 1.0;
    const double lam = cfg.num("a_over_R", 0.3), beta = cfg.num("beta", 0.0);
    const int nx = cfg.integer("nx", 25), nz = cfg.integer("nz", 57);
    const double z_extent = cfg.num("z_extent", 7.0);  // the grid spans |z| <= z_extent (units of a)
    const int prec = factor_precision(cfg);
    Timer clock;
    // confined: R = a / (a/R); unbounded: R only sets the grid's half-width x_extent / 0.97
    const double R = lam > 0 ? a / lam : cfg.num("x_extent", 4.0) / 0.97;
    std::unique_ptr<Tube> tube;
    if (lam > 0)
        tube = std::make_unique<Tube>(R, cfg.num("L_over_R", 4.0) * R, a, cfg.integer("n_theta", 24),
                                      cfg.integer("n_r", 4), 1.0, prec);
    const double b = lam > 0 ? beta * (R - a) : 0.0;
    const V3 c{b, 0, 0}, ez{0, 0, 1};
    SwimmerSolver s(tube.get(), icosphere(a, cfg.integer("sphere_level", 3), c), c, 1.0, prec);
    const std::vector<V3> slips[2] = {s.slip(ez, 1, 0), s.slip(ez, 0, 1)};
    const auto modes = s.solve_free({slips[0], slips[1]});

    // grid in the plane y = 0; points within 1.03 a of the centre are masked (NaN rows)
    std::vector<V3> grid, fluid;
    std::vector<bool> masked;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < nz; ++j) {
            const V3 p{-0.97 * R + 1.94 * R * i / (nx - 1), 0, -z_extent + 2 * z_extent * j / (nz - 1)};
            const bool in = norm(p - c) < 1.03 * a;
            grid.push_back(p);
            masked.push_back(in);
            if (!in) fluid.push_back(p);
        }
    for (int m = 0; m < 2; ++m) {
        const std::vector<V3> vel = s.velocity(modes[m], slips[m], fluid);
        const V3& U = modes[m].U;
        Csv f(out + (m == 0 ? "/field_B1.csv" : "/field_B2.csv"), "x,z,ux,uy,uz",
              {"lab-frame velocity in the plane y=0 for B" + std::to_string(m + 1) +
                   "=1; points inside the sphere are NaN. a_over_R=" + std::to_string(lam) +
                   " beta=" + std::to_string(beta),
               "swimmer U=(" + std::to_string(U.x) + "," + std::to_string(U.y) + "," + std::to_string(U.z) + ")"});
        size_t next = 0;
        for (size_t k = 0; k < grid.size(); ++k) {
            if (masked[k]) {
                f.row({grid[k].x, grid[k].z, NAN, NAN, NAN});
            } else {
                const V3& v = vel[next++];
                f.row({grid[k].x, grid[k].z, v.x, v.y, v.z});
            }
        }
    }
    Csv info(out + "/field_info.csv", "a_over_R,beta,R,UB1z,UB2z,OB1y");
    info.row({lam, beta, lam > 0 ? R : 0.0, modes[0].U.z, modes[1].U.z, modes[0].Om.y});
    std::printf("[field] %zu fluid points, 2 modes, U_B1 = %.5f: %.2f s\n", fluid.size(), modes[0].U.z,
                clock.seconds());
}

// ---------------------------------------------------------------- trajectory
void trajectory_mode(const Config& cfg, const std::string& out) {
    TrajectoryParams P;  // defaults of src/trajectory.hpp:23-36
    P.alpha = cfg.num("alpha", P.alpha);
    P.a_over_R = cfg.num("a_over_R", P.a_over_R);
    P.beta0 = cfg.num("beta0", P.beta0);
    P.pitch0_deg = cfg.num("pitch0_deg", P.pitch0_deg);
    P.yaw0_deg = cfg.num("yaw0_deg", P.yaw0_deg);
    P.dt = cfg.num("dt", P.dt);
    P.t_max = cfg.num("t_max", P.t_max);
    P.L_over_R = cfg.num("L_over_R", P.L_over_R);
    P.n_theta = cfg.integer("n_theta", P.n_theta);
    P.sphere_level = cfg.integer("sphere_level", P.sphere_level);
    P.beta_stop = cfg.num("beta_stop", P.beta_stop);
    P.budget_s = cfg.num("budget_s", P.budget_s);
    P.n_r = cfg.integer("n_r", P.n_r);
    P.precision = factor_precision(cfg);
    ensure_dir(out);
    run_trajectory(P, out + "/" + cfg.str("name", "trajectory"));
}

using ModeFn = std::function<void(const Config&, const std::string&)>;
const std::pair<const char*, ModeFn> kModes[] = {
    {"duct", duct_mode}, {"kinematics", kinematics_mode}, {"field", field_mode}, {"trajectory", trajectory_mode}};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <input-file> [key=value ...]\n", argv[0]);
        return 1;
    }
    try {
        Config cfg = Config::load(argv[1]);
        for (int i = 2; i < argc; ++i) cfg.parse_line(argv[i]);
        const std::string mode = cfg.str("mode");
        const std::string out = cfg.str("output_dir", "outputs/" + mode);
        ensure_dir(out);
        for (const auto& m : kModes)
            if (mode == m.first) {
                m.second(cfg, out);
                return 0;
            }
        throw std::runtime_error("unknown mode '" + mode + "' (duct|kinematics|field|trajectory)");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
