// Option parsing for meta::Inputparameters: "--name=value" with the reference's
// option names (meta/parameters*.cpp); unknown names are an error, like
// boost::program_options in the reference.
#include "../include/meta/inputparameters.hpp"

#include <cctype>
#include <cstdlib>
#include <functional>
#include <map>
#include <stdexcept>

namespace {

bool parse_bool(const std::string& v)
{
    if (v == "true" || v == "1" || v == "on" || v == "yes") return true;
    if (v == "false" || v == "0" || v == "off" || v == "no") return false;
    throw std::invalid_argument("not a boolean: " + v);
}
double parse_double(const std::string& v)
{
    size_t pos = 0;
    double d   = std::stod(v, &pos);
    if (pos != v.size()) throw std::invalid_argument("not a number: " + v);
    return d;
}
std::string lower(std::string v)
{
    for (auto& ch : v) ch = (char)std::tolower((unsigned char)ch);
    return v;
}
int parse_int(const std::string& v)
{
    size_t pos = 0;
    int i      = std::stoi(v, &pos);
    if (pos != v.size()) throw std::invalid_argument("not an integer: " + v);
    return i;
}

}  // namespace

namespace meta {

Inputparameters::Inputparameters(int argc, const char** argv)
{
    std::vector<std::string> a;
    for (int i = 1; i < argc; ++i) a.emplace_back(argv[i]);  // argv[0] is the program name
    parse(a);
}

Inputparameters::Inputparameters(const std::vector<std::string>& args)
{
    parse(args);
}

void Inputparameters::parse(const std::vector<std::string>& args)
{
    std::map<std::string, std::function<void(const std::string&)>> opt = {
        {"nSpace", [&](const std::string& v) { nspace = parse_int(v); }},
        {"nTime", [&](const std::string& v) { ntime = parse_int(v); }},
        {"precision", [&](const std::string& v) { precision = parse_int(v); }},
        {"deviceId", [&](const std::string& v) { device_id = parse_int(v); }},
        {"nDevices", [&](const std::string& v) {
             if (parse_int(v) != 1) throw std::invalid_argument("this build drives one GPU: nDevices must be 1");
         }},
        {"useGPU", [&](const std::string& v) { parse_bool(v); }},
        {"useReconstruct12", [&](const std::string& v) { use_rec12 = parse_bool(v); }},
        {"startCondition", [&](const std::string& v) {
             if (v == "cold_start" || v == "cold")
                 startcondition = common::cold_start;
             else if (v == "hot_start" || v == "hot")
                 startcondition = common::hot_start;
             else if (v == "continue" || v == "start_from_source")
                 startcondition = common::start_from_source;
             else
                 throw std::invalid_argument("unknown start condition " + v);
         }},
        {"initialConf", [&](const std::string& v) { initial_conf = v; }},
        {"hostSeed", [&](const std::string& v) { host_seed = (uint32_t)parse_int(v); }},
        {"fermionAction", [&](const std::string& v) {
             if (v == "wilson")
                 fermact = common::wilson;
             else if (v == "twistedmass")
                 fermact = common::twistedmass;
             else if (v == "rooted_stagg")
                 fermact = common::rooted_stagg;
             else
                 throw std::invalid_argument("Unkown fermion action! wilson, twistedmass, rooted_stagg: " + v);
         }},
        {"beta", [&](const std::string& v) { beta = parse_double(v); }},
        {"nSources", [&](const std::string& v) { num_sources = parse_int(v); }},
        {"sourceX", [&](const std::string& v) { source_x = parse_int(v); }},
        {"sourceY", [&](const std::string& v) { source_y = parse_int(v); }},
        {"sourceZ", [&](const std::string& v) { source_z = parse_int(v); }},
        {"sourceT", [&](const std::string& v) { source_t = parse_int(v); }},
        {"sourceType", [&](const std::string& v) {  // meta/parametersSources.cpp:93-108
             const std::string l = lower(v);
             if (l == "point") sourcetype = common::point;
             else if (l == "volume") sourcetype = common::volume;
             else if (l == "timeslice") sourcetype = common::timeslice;
             else if (l == "zslice") sourcetype = common::zslice;
             else throw std::invalid_argument("sourceType: point, volume, timeslice or zslice: " + v);
         }},
        {"sourceContent", [&](const std::string& v) {  // meta/parametersSources.cpp:110-125
             const std::string l = lower(v);
             if (l == "one") sourcecontent = common::one;
             else if (l == "z4") sourcecontent = common::z4;
             else if (l == "gaussian") sourcecontent = common::gaussian;
             else if (l == "z2") sourcecontent = common::z2;
             else throw std::invalid_argument("sourceContent: one, z4, gaussian or z2: " + v);
         }},
        {"measureCorrelators", [&](const std::string& v) { parse_bool(v); }},
        {"correlatorDirection", [&](const std::string& v) {
             corr_dir = parse_int(v);
             if (corr_dir != 0 && corr_dir != 3)
                 throw std::invalid_argument("correlatorDirection must be 0 (t) or 3 (z): " + v);
         }},
        {"kappa", [&](const std::string& v) { kappa = parse_double(v); }},
        {"mu", [&](const std::string& v) { mu = parse_double(v); }},
        {"mass", [&](const std::string& v) { mass = parse_double(v); }},
        {"thetaFermionSpatial", [&](const std::string& v) { theta_s = parse_double(v); }},
        {"thetaFermionTemporal", [&](const std::string& v) { theta_t = parse_double(v); }},
        {"useChemicalPotentialRe", [&](const std::string& v) { use_cp_re = parse_bool(v); }},
        {"useChemicalPotentialIm", [&](const std::string& v) { use_cp_im = parse_bool(v); }},
        {"chemicalPotentialRe", [&](const std::string& v) { cp_re = parse_double(v); }},
        {"chemicalPotentialIm", [&](const std::string& v) { cp_im = parse_double(v); }},
        {"useEO", [&](const std::string& v) { use_eo = parse_bool(v); }},
        {"useKernelMergingSpinor", [&](const std::string& v) { merge_spinor = parse_bool(v); }},
        {"useKernelMergingFermionMatrix", [&](const std::string& v) { merge_fermion = parse_bool(v); }},
        {"solver", [&](const std::string& v) {
             if (v == "cg")
                 solver = common::cg;
             else if (v == "bicgstab")
                 solver = common::bicgstab;
             else if (v == "bicgstab_save")
                 solver = common::bicgstab_save;
             else
                 throw std::invalid_argument("unknown solver " + v);
         }},
        {"solverMaxIterations", [&](const std::string& v) { cgmax = parse_int(v); }},
        {"solverMinIterations", [&](const std::string& v) { cg_min_iter = parse_int(v); }},
        {"solverPrecision", [&](const std::string& v) { solver_prec = parse_double(v); }},
        {"solverRestartEvery", [&](const std::string& v) { iter_refresh = parse_int(v); }},
        {"solverResiduumCheckEvery", [&](const std::string& v) { cg_block = parse_int(v); }},
        {"solverUseAsyncCopy", [&](const std::string& v) { async_copy = parse_bool(v); }},
    };
    for (const std::string& a : args) {
        if (a.rfind("--", 0) != 0) throw std::invalid_argument("expected --name=value, got " + a);
        const size_t eq        = a.find('=');
        const std::string name = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
        const std::string val  = eq == std::string::npos ? "true" : a.substr(eq + 1);
        auto it                = opt.find(name);
        if (it == opt.end()) throw std::invalid_argument("unknown option --" + name);
        it->second(val);
    }
    if (nspace < 2 || ntime < 2 || nspace % 2 || ntime % 2)
        throw std::invalid_argument("nSpace and nTime must be even and >= 2");
    if (precision != 64) throw std::invalid_argument("the drop-in operators run in FP64 (precision=64)");
    if (iter_refresh < 1 || cg_block < 1 || cgmax < 0) throw std::invalid_argument("bad solver iteration settings");
}

double get_mubar(const Inputparameters& params)
{
    return 2. * params.get_kappa() * params.get_mu();
}

size_t get_vol4d(const Inputparameters& params)
{
    const size_t ns = params.get_nspace();
    return ns * ns * ns * params.get_ntime();
}

}  // namespace meta
