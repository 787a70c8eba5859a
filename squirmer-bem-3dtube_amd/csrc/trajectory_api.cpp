// trajectory_api.cpp -- C entry point of the trajectory integrator (include/trajectory.hpp
// compiled into libsqbem, so C, Python and C++ callers share one implementation).
#include <cstring>
#include <string>

#include "../../include/trajectory.hpp"

namespace sq {
void set_last_error(const std::string& msg);
}

extern "C" int sqbem_trajectory_run(double alpha, double a_over_R, double beta0, double pitch0_deg,
                                    double yaw0_deg, double dt, double t_max, double L_over_R, int n_theta,
                                    int sphere_level, double beta_stop, double budget_s, int n_r, int precision,
                                    const char* out_prefix, int verbose, char* status, int status_len) {
    try {
        if (!out_prefix) throw std::runtime_error("sqbem_trajectory_run: null output prefix");
        bem::TrajectoryParams P;
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
        P.n_r = n_r;
        P.precision = precision;
        const std::string st = bem::run_trajectory(P, out_prefix, verbose != 0);
        if (status && status_len > 0) {
            std::strncpy(status, st.c_str(), (size_t)status_len - 1);
            status[status_len - 1] = 0;
        }
        return SQBEM_OK;
    } catch (const std::exception& e) {
        sq::set_last_error(e.what());
        return SQBEM_ERR_ARG;
    }
}

extern "C" int sqbem_trajectory_gradient(double alpha, double a_over_R, double beta0, double pitch0_deg,
                                         double yaw0_deg, double dt, int n_steps, int ckpt_every, double L_over_R,
                                         int n_theta, int sphere_level, int n_r, int adjoint, int verbose,
                                         double* out) {
    try {
        if (!out) throw std::runtime_error("sqbem_trajectory_gradient: null output");
        bem::TrajectoryParams P;
        P.alpha = alpha;
        P.a_over_R = a_over_R;
        P.beta0 = beta0;
        P.pitch0_deg = pitch0_deg;
        P.yaw0_deg = yaw0_deg;
        P.dt = dt;
        P.L_over_R = L_over_R;
        P.n_theta = n_theta;
        P.sphere_level = sphere_level;
        P.n_r = n_r;
        P.precision = SQBEM_FP64;
        const bem::TrajectoryGradient G = bem::trajectory_gradient(P, n_steps, ckpt_every, adjoint != 0, verbose != 0);
        out[0] = G.beta_final;
        out[1] = G.tilt_final_deg;
        out[2] = G.dbeta_dbeta0;
        out[3] = G.dbeta_dpitch0;
        out[4] = G.dtilt_dbeta0;
        out[5] = G.dtilt_dpitch0;
        out[6] = G.evaluations;
        return SQBEM_OK;
    } catch (const std::exception& e) {
        sq::set_last_error(e.what());
        return SQBEM_ERR_ARG;
    }
}
