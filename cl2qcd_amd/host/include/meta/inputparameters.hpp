// Run parameters of the inversion path (meta/inputparameters.cpp and the option
// groups of meta/parameters*.cpp), read from "--name=value" arguments with the
// reference's option names and defaults.  Unknown options are rejected.
#pragma once

#include <string>
#include <vector>

#include "../common_types.hpp"

namespace meta {

class Inputparameters {
  public:
    Inputparameters(int argc, const char** argv);
    explicit Inputparameters(const std::vector<std::string>& args);

    int get_nspace() const noexcept { return nspace; }
    int get_ntime() const noexcept { return ntime; }
    int get_precision() const noexcept { return precision; }
    int get_device_id() const noexcept { return device_id; }
    bool get_use_rec12() const noexcept { return use_rec12; }
    common::startcondition get_startcondition() const noexcept { return startcondition; }
    const std::string& get_sourcefile() const noexcept { return initial_conf; }
    int get_num_sources() const noexcept { return num_sources; }
    int get_source_x() const noexcept { return source_x; }
    int get_source_y() const noexcept { return source_y; }
    int get_source_z() const noexcept { return source_z; }
    int get_source_t() const noexcept { return source_t; }
    int get_corr_dir() const noexcept { return corr_dir; }
    common::sourcetypes get_sourcetype() const noexcept { return sourcetype; }
    common::sourcecontents get_sourcecontent() const noexcept { return sourcecontent; }
    uint32_t get_host_seed() const noexcept { return host_seed; }

    common::action get_fermact() const noexcept { return fermact; }
    double get_beta() const noexcept { return beta; }
    double get_kappa() const noexcept { return kappa; }
    double get_mu() const noexcept { return mu; }
    double get_mass() const noexcept { return mass; }
    double get_theta_fermion_spatial() const noexcept { return theta_s; }
    double get_theta_fermion_temporal() const noexcept { return theta_t; }
    bool get_use_chem_pot_re() const noexcept { return use_cp_re; }
    bool get_use_chem_pot_im() const noexcept { return use_cp_im; }
    double get_chem_pot_re() const noexcept { return cp_re; }
    double get_chem_pot_im() const noexcept { return cp_im; }
    bool get_use_eo() const noexcept { return use_eo; }
    bool get_use_merge_kernels_fermion() const noexcept { return merge_fermion; }
    bool get_use_merge_kernels_spinor() const noexcept { return merge_spinor; }

    common::solver get_solver() const noexcept { return solver; }
    int get_cgmax() const noexcept { return cgmax; }
    int get_cg_minimum_iteration_count() const noexcept { return cg_min_iter; }
    double get_solver_prec() const noexcept { return solver_prec; }
    int get_iter_refresh() const noexcept { return iter_refresh; }
    int get_cg_iteration_block_size() const noexcept { return cg_block; }
    bool get_cg_use_async_copy() const noexcept { return async_copy; }

  private:
    void parse(const std::vector<std::string>& args);

    int nspace = 4, ntime = 8, precision = 64, device_id = 0;
    bool use_rec12 = false;
    common::startcondition startcondition = common::cold_start;
    std::string initial_conf;
    int num_sources = 12, source_x = 0, source_y = 0, source_z = 0, source_t = 0;
    int corr_dir = 3;  // meta/parametersObs.cpp:72: z by default
    common::sourcetypes sourcetype       = common::point;  // meta/parametersSources.cpp:67-79
    common::sourcecontents sourcecontent = common::one;
    uint32_t host_seed = 4815;
    common::action fermact = common::wilson;
    double beta = 4.0;
    double kappa = 0.125, mu = 0.006, mass = 0.1, theta_s = 0., theta_t = 0.;
    bool use_cp_re = false, use_cp_im = false;
    double cp_re = 0., cp_im = 0.;
    bool use_eo = true, merge_fermion = false, merge_spinor = false;
    common::solver solver = common::bicgstab;
    int cgmax = 1000, cg_min_iter = 0, iter_refresh = 100, cg_block = 10;
    double solver_prec = 1e-23;
    bool async_copy = false;
};

/* mubar = 2 kappa mu (meta/util.cpp:56-58) */
double get_mubar(const Inputparameters& params);
size_t get_vol4d(const Inputparameters& params);

}  // namespace meta
