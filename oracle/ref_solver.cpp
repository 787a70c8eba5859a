// ref_solver.cpp -- C entry points around the reference's own solver headers
// (src/swimmer.hpp, src/duct_flow.hpp, src/trajectory.hpp over src/utils/*.hpp), compiled
// from the reference sources where they lie into oracle/_ref/libref_solver.so and linked
// against the LP64 OpenBLAS this image ships (the four LAPACK/BLAS names the reference
// declares, src/utils/linalg.hpp:17-26, are mapped onto that library's exported names at
// compile time by oracle/Makefile; nothing is stubbed).
//
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_ref_solver.py to write 17-digit
// fixtures, by tests/ to pin the oracle restatement, and by bench.py's CPU baseline, which
// times the reference's own FP64 solve on the host cores. Never linked into the product.
//
// The reference's Tube constructor assembles and factors in one private call
// (src/swimmer.hpp:46-60). ref_tube_split builds the same object in two timed halves: the
// assembly loop of Tube::factor (src/swimmer.hpp:48-57: panel_integrals per pair, SL/(8 pi
// mu), Tube::deflate) and LUFactor::factor, filling the public members the constructor
// fills; the arithmetic is the same calls in the same order, so the factors are identical.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "duct_flow.hpp"
#include "swimmer.hpp"
#include "trajectory.hpp"

extern "C" {
void scipy_openblas_set_num_threads(int);
int scipy_openblas_get_num_threads(void);
char* scipy_openblas_get_config(void);
}

using namespace bem;

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

std::vector<V3> to_v3(const double* a, int n) {
    std::vector<V3> v(n);
    for (int i = 0; i < n; ++i) v[i] = V3{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    return v;
}
void from_v3(const std::vector<V3>& v, double* a) {
    for (size_t i = 0; i < v.size(); ++i)
        for (int c = 0; c < 3; ++c) a[3 * i + c] = v[i][c];
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }
void ref_blas_threads(int n) { scipy_openblas_set_num_threads(n); }
int ref_blas_get_threads() { return scipy_openblas_get_num_threads(); }
const char* ref_blas_config() { return scipy_openblas_get_config(); }

// Tube(R, L, a, n_theta, n_r) as the reference constructs it (src/swimmer.hpp:40-45)
void* ref_tube(double R, double L, double a, int nth, int nr) {
    try {
        return new Tube(R, L, a, nth, nr);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// The same Tube in two timed halves (see the file comment). z == nullptr: the graded
// swimmer_tube_mesh(R, L, a, n_theta, n_r); else tube_mesh(R, z[0..nz), n_theta, n_r).
void* ref_tube_split(double R, double L, double a, int nth, int nr, const double* z, int nz, double* t_assembly,
                     double* t_factor) {
    try {
        Timer clock;
        auto* t = new Tube();
        t->R = R;
        t->L = L;
        t->mu = 1.0;
        t->mesh = z ? tube_mesh(R, std::vector<double>(z, z + nz), nth, nr) : swimmer_tube_mesh(R, L, a, nth, nr);
        t->Nt = t->mesh.P();
        const int Nt = t->Nt, n = 3 * Nt;
        Mat T(n, n);
        double SL[9], DL[9];
        for (int i = 0; i < Nt; ++i)
            for (int p = 0; p < Nt; ++p) {
                panel_integrals(t->mesh, p, t->mesh.centroid[i], i == p, SL, DL);
                for (int c = 0; c < 3; ++c)
                    for (int b = 0; b < 3; ++b) T(3 * i + c, 3 * p + b) = SL[3 * c + b] / (EIGHT_PI * t->mu);
            }
        Tube::deflate(t->mesh, 0, Nt, T, 0, t->mu);
        const double t1 = clock.seconds();
        t->lu.factor(std::move(T));
        if (t_assembly) *t_assembly = t1;
        if (t_factor) *t_factor = clock.seconds() - t1;
        return t;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
int ref_tube_np(void* t) { return ((Tube*)t)->Nt; }
void ref_tube_free(void* t) { delete (Tube*)t; }

// SwimmerSolver(tube or unbounded, icosphere(a, level, c), c) (src/swimmer.hpp:92-97)
void* ref_swimmer(void* tube, double a, int level, const double* c) {
    try {
        const V3 xc{c[0], c[1], c[2]};
        return new SwimmerSolver((const Tube*)tube, icosphere(a, level, xc), xc);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
int ref_swimmer_np(void* s) { return ((SwimmerSolver*)s)->Ns; }
void ref_swimmer_times(void* s, double* t_assembly, double* t_schur) {
    *t_assembly = ((SwimmerSolver*)s)->t_assembly;
    *t_schur = ((SwimmerSolver*)s)->t_schur;
}
void ref_swimmer_free(void* s) { delete (SwimmerSolver*)s; }

// solve_free for nslip slips slip(e, B[2k], B[2k+1]) (src/swimmer.hpp:100-139); U, Om
// [3 nslip], fs [nslip 3 Ns], ft [nslip 3 Nt] (may be null), *t_solve seconds
int ref_swimmer_solve_free(void* sp, const double* e, int nslip, const double* B, double* U, double* Om, double* fs,
                           double* ft, double* t_solve) {
    return guard([&] {
        auto* s = (SwimmerSolver*)sp;
        const V3 ev{e[0], e[1], e[2]};
        std::vector<std::vector<V3>> slips;
        for (int k = 0; k < nslip; ++k) slips.push_back(s->slip(ev, B[2 * k], B[2 * k + 1]));
        Timer clock;
        const auto sol = s->solve_free(slips);
        if (t_solve) *t_solve = clock.seconds();
        for (int k = 0; k < nslip; ++k) {
            for (int c = 0; c < 3; ++c) {
                U[3 * k + c] = sol[k].U[c];
                Om[3 * k + c] = sol[k].Om[c];
            }
            if (fs) from_v3(sol[k].f_sphere, fs + (size_t)k * 3 * s->Ns);
            if (ft && s->Nt) from_v3(sol[k].f_tube, ft + (size_t)k * 3 * s->Nt);
        }
    });
}

// solve_prescribed + force (src/swimmer.hpp:142-160)
int ref_swimmer_towed(void* sp, const double* U, const double* Om, double* F, double* fs, double* ft) {
    return guard([&] {
        auto* s = (SwimmerSolver*)sp;
        const auto sol = s->solve_prescribed(V3{U[0], U[1], U[2]}, V3{Om[0], Om[1], Om[2]});
        const V3 f = s->force(sol);
        for (int c = 0; c < 3; ++c) F[c] = f[c];
        if (fs) from_v3(sol.f_sphere, fs);
        if (ft && s->Nt) from_v3(sol.f_tube, ft);
    });
}

// SwimmerSolver::velocity (src/swimmer.hpp:163-173) of the solution (U, Om, fs, ft) with the
// slip field slip(e, B1, B2) at npts points
int ref_swimmer_velocity(void* sp, const double* U, const double* Om, const double* fs, const double* ft,
                         const double* e, double B1, double B2, const double* pts, int npts, double* out) {
    return guard([&] {
        auto* s = (SwimmerSolver*)sp;
        SwimmerSolution sol;
        sol.U = V3{U[0], U[1], U[2]};
        sol.Om = V3{Om[0], Om[1], Om[2]};
        sol.f_sphere = to_v3(fs, s->Ns);
        if (s->Nt) sol.f_tube = to_v3(ft, s->Nt);
        const auto v = s->velocity(sol, s->slip(V3{e[0], e[1], e[2]}, B1, B2), to_v3(pts, npts));
        from_v3(v, out);
    });
}

// run_trajectory (src/trajectory.hpp:42-151) with the given parameters; status receives
// "finished", "wall contact" or "paused"
int ref_trajectory(double alpha, double a_over_R, double beta0, double pitch0_deg, double yaw0_deg, double dt,
                   double t_max, double L_over_R, int n_theta, int sphere_level, double beta_stop, double budget_s,
                   const char* out_prefix, int verbose, char* status, int status_len) {
    return guard([&] {
        TrajectoryParams P;
        P.alpha = alpha;
        P.a_over_R = a_over_R;
        P.beta0 = beta0;
        P.pitch0_deg = pitch0_deg;
        P.yaw0_deg = yaw0_deg;
        P.dt = dt;
        P.t_max = t_max;
        P.L_over_R = L_over_R;
        P.n_theta = n_theta;
        P.sphere_level = sphere_level;
        P.beta_stop = beta_stop;
        P.budget_s = budget_s;
        const std::string st = run_trajectory(P, out_prefix, verbose != 0);
        std::snprintf(status, status_len, "%s", st.c_str());
    });
}

// DuctFlow(R, L, n_theta, n_z, n_r, mu).solve(p_in, p_out) (src/duct_flow.hpp:32-63); u, f
// [3 P], *Q; np receives P
int ref_duct(double R, double L, int nth, int nz, int nr, double mu, double p_in, double p_out, double* u, double* f,
             double* Q, int* np) {
    return guard([&] {
        DuctFlow d(R, L, nth, nz, nr, mu);
        const auto s = d.solve(p_in, p_out);
        *np = d.mesh.P();
        if (u) from_v3(s.u, u);
        if (f) from_v3(s.f, f);
        *Q = s.Q;
    });
}

}  // extern "C"
