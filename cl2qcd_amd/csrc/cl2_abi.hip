// C ABI of the gfx950 fermion-inversion kernels (declared in include/cl2qcd.h).
// Argument validation happens here, on the host, before any launch: shapes,
// parities, formats and pointer aliasing that a kernel cannot survive are
// rejected with CL2_E_ARG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <initializer_list>
#include <vector>

#include "cl2_internal.hpp"

using namespace cl2;

namespace {

std::string g_create_error;
std::mutex g_create_mutex;

/* Reduction scratch of cl2_scalar_product_eo_noalloc: one double2 partial sum per stage-1 block, at most
 * BLAS_MAX_BLOCKS of them (1 MiB).  Every device the library runs on holds its own copy until the process
 * ends; host code reaches it only through the per-device address that hipGetSymbolAddress returns. */
constexpr unsigned NOALLOC_PARTIALS = BLAS_MAX_BLOCKS;
__attribute__((used)) static __device__ __align__(256) double scratch[2 * NOALLOC_PARTIALS];
constexpr int NOALLOC_MAX_DEVICES = 64;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

/* the device address of scratch on the current device (which is `device`), looked up once per device */
hipError_t scratch_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    double* ptr     = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (ptr == nullptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

double* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

int fail(cl2_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define CL2_HIP(ctx, call)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(ctx, CL2_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CL2_REQUIRE(ctx, cond, msg) \
    do {                            \
        if (!(cond)) return fail(ctx, CL2_E_ARG, msg); \
    } while (0)

size_t spinor_bytes(const cl2_ctx* c, int prec)
{
    if (prec == CL2_F16) return (size_t)52 * c->g.veo;  // 12 complex halves + one float norm per site
    return (size_t)12 * c->g.veo * (prec == CL2_F64 ? 16 : 8);
}
size_t su3vec_bytes(const cl2_ctx* c, int prec)
{
    return (size_t)3 * c->g.veo * (prec == CL2_F64 ? 16 : 8);
}
size_t gauge_bytes(const cl2_ctx* c, int fmt)
{
    // 4 dirs x 2 parities x Veo links, Veo padded to whole tiles of 64 sites
    const size_t tiled = (size_t)8 * link_sites(c->g.veo);
    switch (fmt) {
        case CL2_GAUGE_F64_18: return tiled * 9 * 16;
        case CL2_GAUGE_F64_12: return tiled * 6 * 16;
        case CL2_GAUGE_F32_12: return tiled * 3 * 16;
        case CL2_GAUGE_F16_12: return tiled * 24;
        default: return 0;
    }
}
/* precisions of every field kind */
bool valid_prec(int p)
{
    return p == CL2_F64 || p == CL2_F32;
}
/* precisions of Wilson eo spinors (FP16 with per-site scale included) */
bool valid_wprec(int p)
{
    return valid_prec(p) || p == CL2_F16;
}
/* FP32 Wilson spinors are read and written as 16-byte component pairs */
static bool aligned16(const void* p)
{
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}
static bool all_aligned16(std::initializer_list<const void*> ps)
{
    for (const void* p : ps)
        if (!aligned16(p)) return false;
    return true;
}
#define CL2_PAIRS_ALIGNED(c, prec, ...)                                                                  \
    CL2_REQUIRE(c, (prec) == CL2_F64 || all_aligned16({__VA_ARGS__}),                                     \
                "FP32 and FP16 Wilson spinor fields must be 16-byte aligned")

bool valid_fmt(int f)
{
    return f == CL2_GAUGE_F64_18 || f == CL2_GAUGE_F64_12 || f == CL2_GAUGE_F32_12 || f == CL2_GAUGE_F16_12;
}
int fmt_prec(int f)
{
    return f == CL2_GAUGE_F16_12 ? CL2_F16 : (f == CL2_GAUGE_F32_12 ? CL2_F32 : CL2_F64);
}

inline int site_parity(const cl2_ctx* c, size_t s)
{
    const size_t ns = c->g.ns;
    size_t x = s % ns, r = s / ns;
    size_t y = r % ns;
    r /= ns;
    size_t z = r % ns, t = r / ns;
    return (int)((x + y + z + t) & 1);
}

/* a temporary device buffer freed on scope exit (import / export staging only) */
struct TempDev {
    void* p = nullptr;
    ~TempDev()
    {
        if (p) (void)hipFree(p);
    }
};

cl2_scalar value(double re, double im = 0.)
{
    cl2_scalar s;
    s.kind  = CL2_SCALAR_VALUE;
    s.re    = re;
    s.im    = im;
    s.dev   = nullptr;
    s.index = 0;
    return s;
}

int check_scalar(cl2_ctx* c, const cl2_scalar* s)
{
    CL2_REQUIRE(c, s != nullptr, "scalar argument is null");
    CL2_REQUIRE(c, s->kind >= CL2_SCALAR_VALUE && s->kind <= CL2_SCALAR_DEV_VECTOR, "unknown scalar kind");
    CL2_REQUIRE(c, s->kind == CL2_SCALAR_VALUE || s->dev != nullptr, "device scalar pointer is null");
    CL2_REQUIRE(c, s->kind != CL2_SCALAR_DEV_VECTOR || s->index >= 0, "negative vector index");
    return CL2_OK;
}

/* elementwise BLAS dispatch shared by the Wilson and staggered entry points */
int run_blas(cl2_ctx* c, int mode, size_t n, void* out, const void* x, const void* y, const void* z, int prec,
             const cl2_scalar& a, const cl2_scalar& b, double* norm_res, size_t tail, int cg_norm = 0,
             void* out2 = nullptr)
{
    const bool half = prec == CL2_F16;
    CL2_REQUIRE(c, valid_prec(prec) || (half && n == (size_t)12 * c->g.veo),
                half ? "FP16 storage is for Wilson eo spinors only" : "unknown precision");
    CL2_REQUIRE(c, out && x, "null field pointer");
    BlasArgs args;
    args.n             = n;
    args.tail          = tail;
    args.out           = out;
    args.out2          = out2;
    args.x             = x;
    args.y             = y;
    args.z             = z;
    args.a             = a;
    args.b             = b;
    args.mode          = mode;
    args.norm_partials = norm_res ? c->partials : nullptr;
    args.cg_norm       = cg_norm;
    unsigned nb        = 0;
    if (half) {
        args.n = c->g.veo;
        CL2_HIP(c, launch_blas_h(c, args, &nb));
    } else {
        CL2_HIP(c, launch_blas(c, args, prec, &nb));
    }
    if (norm_res) CL2_HIP(c, launch_reduce_stage2(c, nb, norm_res, cg_norm ? 3 : 1, nullptr, 0));
    return CL2_OK;
}

int run_reduce(cl2_ctx* c, double* res, const void* x, const void* y, size_t n, int prec, int mode, int out_kind)
{
    const bool half = prec == CL2_F16;
    CL2_REQUIRE(c, valid_prec(prec) || (half && n == (size_t)12 * c->g.veo),
                half ? "FP16 storage is for Wilson eo spinors only" : "unknown precision");
    CL2_REQUIRE(c, res && x && (mode == R_NORM || y), "null pointer");
    unsigned nb = 0;
    if (half)
        CL2_HIP(c, launch_reduce_stage1_h(c, x, y, mode, &nb));
    else
        CL2_HIP(c, launch_reduce_stage1(c, x, y, n, prec, mode, &nb));
    CL2_HIP(c, launch_reduce_stage2(c, nb, res, out_kind, nullptr, 0));
    return CL2_OK;
}

/* per-right-hand-side operator parameters of a multi-RHS launch */
struct RhsParams {
    double kappa[16];
    double f1[16][2], f2[16][2], scale[16];
};

int run_dslash(cl2_ctx* c, void* out, const void* in, const void* x, const void* gauge, int fmt, int evenodd,
               double kappa, int epilogue, const double f1[2], const double f2[2], double scale, int gamma5,
               int nrhs = 1, const RhsParams* pr = nullptr, double* norm_res = nullptr, void* cg_r = nullptr,
               const double* cg_alpha = nullptr)
{
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, supported_nrhs(nrhs), "unsupported number of right-hand sides (1, 2, 4, 8, 12 or 16)");
    CL2_REQUIRE(c, evenodd == CL2_EVEN || evenodd == CL2_ODD, "evenodd must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, (out || cg_r) && in && gauge, "null field pointer");
    CL2_REQUIRE(c, out != in && cg_r != in, "dslash cannot run in place: out aliases in");
    CL2_REQUIRE(c, cg_r == nullptr || cg_r != x, "the CG residual must differ from the diagonal-term field");
    CL2_REQUIRE(c, epilogue != EPI_AEE || x != nullptr, "null diagonal-term field");
    CL2_REQUIRE(c, epilogue != EPI_AEE || x != in, "diagonal-term field must differ from the hopping input");
    if (fmt == CL2_GAUGE_F32_12) {
        CL2_PAIRS_ALIGNED(c, CL2_F32, in);
        CL2_REQUIRE(c, aligned16(out) && aligned16(x) && aligned16(cg_r),
                    "FP32 Wilson spinor fields must be 16-byte aligned");
    }
    DslashArgs a;
    a.in            = in;
    a.out           = out;
    a.gauge         = gauge;
    a.x             = x;
    a.g             = c->g;
    a.out_parity    = (evenodd == CL2_ODD) ? 0 : 1;  // ODD -> even output sites
    a.epilogue      = epilogue;
    a.gamma5        = gamma5;
    a.f1[0]         = f1[0];
    a.f1[1]         = f1[1];
    a.f2[0]         = f2[0];
    a.f2[1]         = f2[1];
    a.scale         = scale;
    a.norm_partials = norm_res ? c->partials : nullptr;
    a.cg_r          = cg_r;
    a.cg_alpha      = cg_alpha;
    a.per_rhs       = (pr != nullptr && nrhs > 1) ? 1 : 0;
    a.w             = wilson_weights(c, a.per_rhs ? 1.0 : kappa);
    for (int r = 0; r < 16; ++r) {
        const bool use   = a.per_rhs && r < nrhs;
        a.rhs_kappa[r]   = use ? pr->kappa[r] : 0.;
        a.rhs_f1[r][0]   = use ? pr->f1[r][0] : 0.;
        a.rhs_f1[r][1]   = use ? pr->f1[r][1] : 0.;
        a.rhs_f2[r][0]   = use ? pr->f2[r][0] : 0.;
        a.rhs_f2[r][1]   = use ? pr->f2[r][1] : 0.;
        a.rhs_scale[r]   = use ? pr->scale[r] : 0.;
    }
    const cl2_hop_tuning t = hop_tuning(c, fmt, nrhs);
    a.block         = nrhs == 1 ? t.block : DSLASH_BLOCK;
    a.occ           = t.occ;
    a.map           = work_map(c, a.block, nrhs, t);
    a.nrhs          = nrhs;
    a.rb            = a.per_rhs ? 1 : t.rb;  // per-right-hand-side operators: one right-hand side per thread
    // twice the sites per staged block only where the FP64 gathers can use 32-bit offsets
    const int lds =
        (t.lds == 2 && fmt != GAUGE_F32_12 &&
         !(c->tune_mrhs_off32 && mrhs_offsets_narrow(a.g.veo, nrhs, 2 * sizeof(double))))
            ? 1
            : t.lds;
    a.lds = lds;
    if (nrhs > 1 && fmt == CL2_GAUGE_F16_12) {  // FP16 blocks: the staged kernel, one right-hand side per thread
        a.rb  = 1;
        a.lds = 1;
    }
    if (nrhs > 1 && a.lds && a.rb <= 1) a.map = work_map(c, mrhs_sites_per_block(nrhs, a.lds), 1, t);
    CL2_REQUIRE(c, cg_r == nullptr || (cg_alpha != nullptr && norm_res != nullptr), "CG update without its scalars");
    // per-right-hand-side norms come from the kernel with per-site staging (one right-hand side per thread)
    const bool mrhs_norms = nrhs > 1 && norm_res != nullptr;
    if (mrhs_norms) {
        a.rb  = 1;
        a.lds = (lds == 2 && fmt != CL2_GAUGE_F16_12) ? 2 : 1;
        a.map = work_map(c, mrhs_sites_per_block(nrhs, a.lds), 1, t);
    }
    CL2_HIP(c, launch_dslash(c, a, fmt));
    // every single-field variant launches blocks_for(Veo, block) blocks, each writing one partial: |out|^2 as a
    // complex (norm, 0), or with the CG update (|r_new|^2, Re <r_new, r_old>, |r_new|^2, 0); the staged
    // multi-RHS kernel writes (|out|^2 or |r_new|^2, 0 or Re <r_new, r_old>) per right-hand side and block
    if (norm_res && !mrhs_norms)
        CL2_HIP(c, launch_reduce_stage2(c, blocks_for(a.g.veo, a.block), norm_res, cg_r ? 3 : 1, nullptr, 0));
    if (mrhs_norms) {
        const int sb = mrhs_sites_per_block(nrhs, a.lds);
        CL2_HIP(c, launch_block_reduce2(c, reinterpret_cast<const double2*>(c->partials),
                                        (unsigned)((a.g.veo + sb - 1) / sb), norm_res, 0, nrhs));
    }
    return CL2_OK;
}

/* factor pairs (spins 0,1 ; spins 2,3) of the twisted-mass site-diagonal terms */
void tm_factors(int variant, double mubar, double f1[2], double f2[2], double* scale)
{
    *scale = 1.;
    switch (variant) {
        case CL2_MTM_SITEDIAGONAL:
            f1[0] = 1., f1[1] = mubar, f2[0] = 1., f2[1] = -mubar;
            break;
        case CL2_MTM_INVERSE_SITEDIAGONAL:
            f1[0] = 1., f1[1] = -mubar, f2[0] = 1., f2[1] = mubar;
            *scale = 1. / (1. + mubar * mubar);
            break;
        case CL2_MTM_SITEDIAGONAL_MINUS:
            f1[0] = 1., f1[1] = -mubar, f2[0] = 1., f2[1] = mubar;
            break;
        case CL2_MTM_INVERSE_SITEDIAGONAL_MINUS:
            f1[0] = 1., f1[1] = mubar, f2[0] = 1., f2[1] = -mubar;
            *scale = 1. / (1. + mubar * mubar);
            break;
        case CL2_MTM_SITEDIAGONAL_AND_GAMMA5:
            f1[0] = 1., f1[1] = mubar, f2[0] = -1., f2[1] = mubar;
            break;
        default:  // CL2_MTM_SITEDIAGONAL_MINUS_AND_GAMMA5
            f1[0] = 1., f1[1] = -mubar, f2[0] = -1., f2[1] = -mubar;
            break;
    }
}

/* Aee family with two fused kernels: tmp = M^-1 D in (Wilson: D in), then
 * out = M_site in - D tmp [gamma5]   (fermionmatrix.cpp:224-591). */
int run_aee(cl2_ctx* c, void* out, const void* in, void* tmp, const void* gauge, int fmt, int action, double kappa,
            double mubar, bool minus, bool gamma5, int nrhs = 1, const double* kappas = nullptr,
            const double* mubars = nullptr, double* norm_res = nullptr, void* cg_r = nullptr,
            const double* cg_alpha = nullptr)
{
    double f1[2], f2[2], s;
    int rc;
    if (kappas) {  // kappa and mubar per right-hand side (kappas, mubars: nrhs values each)
        RhsParams p1, p2;
        const double one[2] = {1., 0.};
        for (int r = 0; r < nrhs; ++r) {
            p1.kappa[r] = p2.kappa[r] = kappas[r];
            double g1[2], g2[2], gs;
            if (action == CL2_ACTION_WILSON) {
                g1[0] = g2[0] = 1., g1[1] = g2[1] = 0., gs = 1.;
            } else {
                tm_factors(minus ? CL2_MTM_INVERSE_SITEDIAGONAL_MINUS : CL2_MTM_INVERSE_SITEDIAGONAL, mubars[r], g1,
                           g2, &gs);
            }
            p1.f1[r][0] = g1[0], p1.f1[r][1] = g1[1], p1.f2[r][0] = g2[0], p1.f2[r][1] = g2[1], p1.scale[r] = gs;
            if (action == CL2_ACTION_WILSON) {
                g1[0] = g2[0] = 1., g1[1] = g2[1] = 0.;
            } else {
                tm_factors(minus ? CL2_MTM_SITEDIAGONAL_MINUS : CL2_MTM_SITEDIAGONAL, mubars[r], g1, g2, &gs);
            }
            p2.f1[r][0] = g1[0], p2.f1[r][1] = g1[1], p2.f2[r][0] = g2[0], p2.f2[r][1] = g2[1], p2.scale[r] = 1.;
        }
        rc = run_dslash(c, tmp, in, nullptr, gauge, fmt, CL2_ODD, 1., EPI_MINV, one, one, 1., 0, nrhs, &p1);
        if (rc) return rc;
        return run_dslash(c, out, tmp, in, gauge, fmt, CL2_EVEN, 1., EPI_AEE, one, one, 1., gamma5 ? 1 : 0, nrhs, &p2,
                          norm_res, cg_r, cg_alpha);
    }
    if (action == CL2_ACTION_WILSON) {
        const double one[2] = {1., 0.};
        rc = run_dslash(c, tmp, in, nullptr, gauge, fmt, CL2_ODD, kappa, EPI_PLAIN, one, one, 1., 0, nrhs);
        if (rc) return rc;
        return run_dslash(c, out, tmp, in, gauge, fmt, CL2_EVEN, kappa, EPI_AEE, one, one, 1., gamma5 ? 1 : 0, nrhs,
                          nullptr, norm_res, cg_r, cg_alpha);
    }
    tm_factors(minus ? CL2_MTM_INVERSE_SITEDIAGONAL_MINUS : CL2_MTM_INVERSE_SITEDIAGONAL, mubar, f1, f2, &s);
    rc = run_dslash(c, tmp, in, nullptr, gauge, fmt, CL2_ODD, kappa, EPI_MINV, f1, f2, s, 0, nrhs);
    if (rc) return rc;
    tm_factors(minus ? CL2_MTM_SITEDIAGONAL_MINUS : CL2_MTM_SITEDIAGONAL, mubar, f1, f2, &s);
    return run_dslash(c, out, tmp, in, gauge, fmt, CL2_EVEN, kappa, EPI_AEE, f1, f2, 1., gamma5 ? 1 : 0, nrhs, nullptr,
                      norm_res, cg_r, cg_alpha);
}

/* the eo operators on nrhs right-hand sides ([12][Veo][nrhs] fields, tmp = 2 such fields) */
int run_fermionmatrix(cl2_ctx* c, int op, void* out, const void* in, const void* gauge, int fmt, int action,
                      double kappa, double mubar, void* tmp, int nrhs, const double* kappas = nullptr,
                      const double* mubars = nullptr, double* dot_res = nullptr)
{
    CL2_REQUIRE(c, op >= CL2_OP_AEE && op <= CL2_OP_QPLUSQMINUS_EO, "unknown operator");
    CL2_REQUIRE(c, action == CL2_ACTION_WILSON || action == CL2_ACTION_TWISTEDMASS,
                "unknown fermion action: wilson or twistedmass");
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, out && in && gauge && tmp, "null field pointer");
    const size_t fb = spinor_bytes(c, fmt_prec(fmt)) * (size_t)nrhs;
    void* t0        = tmp;
    void* t1        = static_cast<char*>(tmp) + fb;
    CL2_REQUIRE(c, out != t0 && out != t1 && in != t0 && in != t1, "tmp must not alias in or out");
    const double *ks = kappas, *ms = mubars;
    switch (op) {
        case CL2_OP_AEE: return run_aee(c, out, in, t0, gauge, fmt, action, kappa, mubar, false, false, nrhs, ks, ms);
        case CL2_OP_AEE_AND_GAMMA5:
        case CL2_OP_QPLUS_EO:
            return run_aee(c, out, in, t0, gauge, fmt, action, kappa, mubar, false, true, nrhs, ks, ms);
        case CL2_OP_AEE_MINUS:
            return run_aee(c, out, in, t0, gauge, fmt, action, kappa, mubar, true, false, nrhs, ks, ms);
        case CL2_OP_AEE_MINUS_AND_GAMMA5:
        case CL2_OP_QMINUS_EO: return run_aee(c, out, in, t0, gauge, fmt, action, kappa, mubar, true, true, nrhs, ks, ms);
        default: {
            // QplusQminus = Qplus Qminus with Qplus = Qminus^dagger: <in, out> = |Qminus in|^2, the norm of the
            // intermediate field, summed by its producing kernel
            CL2_REQUIRE(c, dot_res == nullptr || (nrhs == 1 && ks == nullptr), "<in, out> is for single fields");
            int rc = run_aee(c, t1, in, t0, gauge, fmt, action, kappa, mubar, true, true, nrhs, ks, ms, dot_res);
            if (rc) return rc;
            return run_aee(c, out, t1, t0, gauge, fmt, action, kappa, mubar, false, true, nrhs, ks, ms);
        }
    }
}

}  // namespace

extern "C" {

int cl2_ctx_create(cl2_ctx** out, int device, const cl2_lattice* lat)
{
    auto create_fail = [&](int code, const std::string& m) {
        std::lock_guard<std::mutex> lk(g_create_mutex);
        g_create_error = m;
        return code;
    };
    if (!out || !lat) return create_fail(CL2_E_ARG, "null argument");
    *out = nullptr;
    if (lat->ns < 2 || lat->nt < 2 || (lat->ns & 1) || (lat->nt & 1))
        return create_fail(CL2_E_ARG, "lattice extents must be even and >= 2");
    const double vol = (double)lat->ns * lat->ns * lat->ns * lat->nt;
    if (vol / 2 * 12 >= 4294967296.0) return create_fail(CL2_E_ARG, "lattice too large for 32-bit site indices");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return create_fail(CL2_E_ARG, "no such HIP device");
    if (hipSetDevice(device) != hipSuccess) return create_fail(CL2_E_HIP, "hipSetDevice failed");
    if (const hipError_t e = scratch_setup(device); e != hipSuccess)
        return create_fail(CL2_E_HIP, std::string("hipGetSymbolAddress of the reduction scratch: ") + hipGetErrorString(e));
    cl2_ctx* c = new cl2_ctx();
    c->device  = device;
    c->lat     = *lat;
    c->g.ns    = (uint32_t)lat->ns;
    c->g.nt    = (uint32_t)lat->nt;
    c->g.vs    = c->g.ns * c->g.ns * c->g.ns;
    c->g.veo   = c->g.vs * c->g.nt / 2;
    const double PI = 3.14159265358979323846;
    c->bs[0] = std::cos(lat->theta_s * PI / lat->ns);
    c->bs[1] = std::sin(lat->theta_s * PI / lat->ns);
    c->bt[0] = std::cos(lat->theta_t * PI / lat->nt);
    c->bt[1] = std::sin(lat->theta_t * PI / lat->nt);
    double er = lat->use_cp_re ? std::exp(lat->cp_re) : 1., mer = lat->use_cp_re ? std::exp(-lat->cp_re) : 1.;
    double cr = lat->use_cp_im ? std::cos(lat->cp_im) : 1., ci = lat->use_cp_im ? std::sin(lat->cp_im) : 0.;
    c->cp_fwd[0] = er * cr;
    c->cp_fwd[1] = er * ci;
    c->cp_bwd[0] = mer * cr;
    c->cp_bwd[1] = mer * ci;
    // one double2 per block: 16384 blocks for each of up to 16 right-hand sides (block kernels), BLAS_MAX_BLOCKS
    // for an elementwise kernel, one per block of the smallest hopping-kernel block (128 threads), or nrhs per
    // block of the staged multi-RHS kernel (nrhs x ceil(Veo / sites per block) <= Veo + 16)
    c->npartials = std::max<size_t>({(size_t)16384 * 16, (size_t)BLAS_MAX_BLOCKS, blocks_for(c->g.veo, 128),
                                     (size_t)c->g.veo + 64});
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->partials, c->npartials * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->dscratch, 64 * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->ptr_table, 128 * sizeof(void*)) != hipSuccess ||
        hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_stop) != hipSuccess) {
        cl2_ctx_destroy(c);
        return create_fail(CL2_E_HIP, "context allocation failed");
    }
    *out = c;
    return CL2_OK;
}

int cl2_ctx_destroy(cl2_ctx* c)
{
    if (!c) return CL2_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->partials) (void)hipFree(c->partials);
    if (c->dscratch) (void)hipFree(c->dscratch);
    if (c->ptr_table) (void)hipFree(c->ptr_table);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CL2_OK;
}

const char* cl2_last_error(const cl2_ctx* c)
{
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

int cl2_sync(cl2_ctx* c)
{
    if (!c) return CL2_E_ARG;
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    return CL2_OK;
}

void* cl2_get_stream(cl2_ctx* c)
{
    return c ? (void*)c->stream : nullptr;
}

long cl2_eo_sites(const cl2_ctx* c)
{
    return c ? (long)c->g.veo : 0;
}

size_t cl2_field_bytes(const cl2_ctx* c, int kind, int prec_or_format)
{
    if (!c) return 0;
    switch (kind) {
        case CL2_FIELD_SPINOR_EO: return valid_wprec(prec_or_format) ? spinor_bytes(c, prec_or_format) : 0;
        case CL2_FIELD_SU3VEC_EO: return valid_prec(prec_or_format) ? su3vec_bytes(c, prec_or_format) : 0;
        case CL2_FIELD_GAUGE: return gauge_bytes(c, prec_or_format);
        default: return 0;
    }
}

int cl2_alloc(cl2_ctx* c, void** dev, size_t bytes)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dev != nullptr, "null output pointer");
    *dev = nullptr;
    CL2_REQUIRE(c, bytes > 0, "zero-byte allocation");
    CL2_HIP(c, hipSetDevice(c->device));
    CL2_HIP(c, hipMalloc(dev, bytes));
    CL2_HIP(c, hipMemsetAsync(*dev, 0, bytes, c->stream));
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    return CL2_OK;
}

int cl2_free(cl2_ctx* c, void* dev)
{
    if (!c) return CL2_E_ARG;
    if (!dev) return CL2_OK;
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    CL2_HIP(c, hipFree(dev));
    return CL2_OK;
}

int cl2_copy(cl2_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dst && src, "null pointer");
    if (dst == src || bytes == 0) return CL2_OK;
    if (bytes % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0)
        CL2_HIP(c, launch_copy16(c, dst, src, bytes / 16));
    else
        CL2_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return CL2_OK;
}

/* complex index of component k of site i in a device field of S sites: component-major, FP32 Wilson spinors
 * (ncomp 12) in component pairs (the device layout of load_spinor) */
static size_t host_cx(int prec, int ncomp, int k, size_t S, size_t i)
{
    if (prec == CL2_F32 && ncomp == 12) return (size_t)(k >> 1) * 2 * S + 2 * i + (size_t)(k & 1);
    return (size_t)k * S + i;
}

static int complex_field_import(cl2_ctx* c, void* dev, const double* host, int prec, int ncomp)
{
    CL2_REQUIRE(c, dev && host, "null pointer");
    CL2_REQUIRE(c, valid_prec(prec), "unknown precision");
    const size_t S = c->g.veo;
    if (prec == CL2_F64) {
        std::vector<double> stage(2 * ncomp * S);
        for (size_t i = 0; i < S; ++i)
            for (int k = 0; k < ncomp; ++k) {
                stage[2 * (k * S + i)]     = host[2 * ncomp * i + 2 * k];
                stage[2 * (k * S + i) + 1] = host[2 * ncomp * i + 2 * k + 1];
            }
        CL2_HIP(c, hipMemcpyAsync(dev, stage.data(), stage.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        CL2_PAIRS_ALIGNED(c, ncomp == 12 ? prec : CL2_F64, dev);
        std::vector<float> stage(2 * ncomp * S);
        for (size_t i = 0; i < S; ++i)
            for (int k = 0; k < ncomp; ++k) {
                const size_t e   = host_cx(prec, ncomp, k, S, i);
                stage[2 * e]     = (float)host[2 * ncomp * i + 2 * k];
                stage[2 * e + 1] = (float)host[2 * ncomp * i + 2 * k + 1];
            }
        CL2_HIP(c, hipMemcpyAsync(dev, stage.data(), stage.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
    }
    return CL2_OK;
}

static int complex_field_export(cl2_ctx* c, double* host, const void* dev, int prec, int ncomp)
{
    CL2_REQUIRE(c, dev && host, "null pointer");
    CL2_REQUIRE(c, valid_prec(prec), "unknown precision");
    const size_t S = c->g.veo;
    if (prec == CL2_F64) {
        std::vector<double> stage(2 * ncomp * S);
        CL2_HIP(c, hipMemcpyAsync(stage.data(), dev, stage.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < S; ++i)
            for (int k = 0; k < ncomp; ++k) {
                host[2 * ncomp * i + 2 * k]     = stage[2 * (k * S + i)];
                host[2 * ncomp * i + 2 * k + 1] = stage[2 * (k * S + i) + 1];
            }
    } else {
        std::vector<float> stage(2 * ncomp * S);
        CL2_HIP(c, hipMemcpyAsync(stage.data(), dev, stage.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < S; ++i)
            for (int k = 0; k < ncomp; ++k) {
                const size_t e                  = host_cx(prec, ncomp, k, S, i);
                host[2 * ncomp * i + 2 * k]     = stage[2 * e];
                host[2 * ncomp * i + 2 * k + 1] = stage[2 * e + 1];
            }
    }
    return CL2_OK;
}

int cl2_spinor_import(cl2_ctx* c, void* dev, const double* host, int prec)
{
    if (!c) return CL2_E_ARG;
    if (prec != CL2_F16) return complex_field_import(c, dev, host, prec, 12);
    CL2_REQUIRE(c, dev && host, "null pointer");
    TempDev t;
    CL2_HIP(c, hipMalloc(&t.p, spinor_bytes(c, CL2_F64)));
    if (int rc = complex_field_import(c, t.p, host, CL2_F64, 12)) return rc;
    CL2_HIP(c, launch_convert_h(c, dev, CL2_F16, t.p, CL2_F64));
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    return CL2_OK;
}
int cl2_spinor_export(cl2_ctx* c, double* host, const void* dev, int prec)
{
    if (!c) return CL2_E_ARG;
    if (prec != CL2_F16) return complex_field_export(c, host, dev, prec, 12);
    CL2_REQUIRE(c, dev && host, "null pointer");
    TempDev t;
    CL2_HIP(c, hipMalloc(&t.p, spinor_bytes(c, CL2_F64)));
    CL2_HIP(c, launch_convert_h(c, t.p, CL2_F64, dev, CL2_F16));
    return complex_field_export(c, host, t.p, CL2_F64, 12);
}
int cl2_su3vec_import(cl2_ctx* c, void* dev, const double* host, int prec)
{
    if (!c) return CL2_E_ARG;
    return complex_field_import(c, dev, host, prec, 3);
}
int cl2_su3vec_export(cl2_ctx* c, double* host, const void* dev, int prec)
{
    if (!c) return CL2_E_ARG;
    return complex_field_export(c, host, dev, prec, 3);
}

int cl2_gauge_import(cl2_ctx* c, void* dev, const double* host, int fmt)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dev && host, "null pointer");
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    if (fmt == CL2_GAUGE_F16_12) {
        TempDev t;
        CL2_HIP(c, hipMalloc(&t.p, gauge_bytes(c, CL2_GAUGE_F64_18)));
        if (int rc = cl2_gauge_import(c, t.p, host, CL2_GAUGE_F64_18)) return rc;
        CL2_HIP(c, launch_gauge_convert(c, dev, CL2_GAUGE_F16_12, t.p));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
        return CL2_OK;
    }
    const size_t S = c->g.veo, V = 2 * S;
    if (fmt == CL2_GAUGE_F32_12) {
        std::vector<float> stage(gauge_bytes(c, fmt) / sizeof(float));
        for (size_t s = 0; s < V; ++s) {
            const int par = site_parity(c, s);
            const size_t eo = s / 2;
            for (int mu = 0; mu < 4; ++mu) {
                const double* u = host + 18 * (mu + 4 * s);
                for (int k = 0; k < 3; ++k) {
                    float* d = &stage[4 * link_elem(mu * 2 + par, 3, k, (uint32_t)eo, (uint32_t)S)];
                    d[0]     = (float)u[4 * k];
                    d[1]     = (float)u[4 * k + 1];
                    d[2]     = (float)u[4 * k + 2];
                    d[3]     = (float)u[4 * k + 3];
                }
            }
        }
        CL2_HIP(c, hipMemcpyAsync(dev, stage.data(), stage.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    } else {
        const int lanes = (fmt == CL2_GAUGE_F64_18) ? 9 : 6;
        std::vector<double> stage(gauge_bytes(c, fmt) / sizeof(double));
        for (size_t s = 0; s < V; ++s) {
            const int par = site_parity(c, s);
            const size_t eo = s / 2;
            for (int mu = 0; mu < 4; ++mu) {
                const double* u = host + 18 * (mu + 4 * s);
                for (int k = 0; k < lanes; ++k) {
                    double* d = &stage[2 * link_elem(mu * 2 + par, lanes, k, (uint32_t)eo, (uint32_t)S)];
                    d[0]      = u[2 * k];
                    d[1]      = u[2 * k + 1];
                }
            }
        }
        CL2_HIP(c, hipMemcpyAsync(dev, stage.data(), stage.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    return CL2_OK;
}

int cl2_gauge_set_random(cl2_ctx* c, void* dev, int fmt, uint64_t seed, double epsilon)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dev != nullptr, "null pointer");
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, epsilon >= 0. && std::isfinite(epsilon), "epsilon must be finite and >= 0");
    CL2_HIP(c, launch_gauge_random(c, dev, fmt, seed, epsilon));
    return CL2_OK;
}

int cl2_gauge_heatbath(cl2_ctx* c, void* gauge, double beta, int n_sweeps, int n_overrelax, uint64_t seed)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, gauge != nullptr, "null pointer");
    CL2_REQUIRE(c, beta > 0. && std::isfinite(beta), "beta must be positive");
    CL2_REQUIRE(c, n_sweeps >= 0 && n_overrelax >= 0, "negative sweep count");
    for (int s = 0; s < n_sweeps; ++s) {
        const uint64_t key = seed * 1000003ull + (uint64_t)s * (n_overrelax + 1);
        CL2_HIP(c, launch_gauge_update(c, gauge, beta, 0, key));
        for (int o = 0; o < n_overrelax; ++o) CL2_HIP(c, launch_gauge_update(c, gauge, beta, 1, key + 1 + o));
    }
    return CL2_OK;
}

int cl2_gauge_overrelax(cl2_ctx* c, void* gauge, int n_steps)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, gauge != nullptr, "null pointer");
    CL2_REQUIRE(c, n_steps >= 0, "negative sweep count");
    for (int o = 0; o < n_steps; ++o) CL2_HIP(c, launch_gauge_update(c, gauge, 1.0, 1, 0));
    return CL2_OK;
}

int cl2_gauge_plaquette(cl2_ctx* c, const void* gauge, double* plaq)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, gauge && plaq, "null pointer");
    CL2_HIP(c, launch_plaquette(c, gauge, c->dscratch));
    double s = 0.;
    CL2_HIP(c, hipMemcpyAsync(&s, c->dscratch, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    *plaq = s / (18. * 2. * c->g.veo);
    return CL2_OK;
}

int cl2_gauge_convert(cl2_ctx* c, void* dst, int dfmt, const void* src)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dst && src && dst != src, "bad pointers");
    CL2_REQUIRE(c, valid_fmt(dfmt), "unknown gauge format");
    CL2_HIP(c, launch_gauge_convert(c, dst, dfmt, src));
    return CL2_OK;
}

int cl2_gauge_export(cl2_ctx* c, double* host, const void* dev, int fmt)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dev && host, "null pointer");
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    if (fmt == CL2_GAUGE_F16_12) return fail(c, CL2_E_UNSUPPORTED, "export of F16_12 links is not supported");
    const size_t S = c->g.veo, V = 2 * S;
    const int lanes = fmt == CL2_GAUGE_F64_18 ? 9 : 6;
    // staging in the device layout of the FP64 format with the same rows (FP32 12-real -> FP64 12-real)
    std::vector<double> stage(gauge_bytes(c, lanes == 9 ? CL2_GAUGE_F64_18 : CL2_GAUGE_F64_12) / sizeof(double));
    if (fmt == CL2_GAUGE_F32_12) {
        std::vector<float> f(gauge_bytes(c, fmt) / sizeof(float));
        CL2_HIP(c, hipMemcpyAsync(f.data(), dev, f.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
        for (uint32_t dp = 0; dp < 8; ++dp)
            for (uint32_t k = 0; k < 3; ++k)
                for (uint32_t eo = 0; eo < (uint32_t)S; ++eo)
                    for (int q = 0; q < 4; ++q)
                        stage[2 * link_elem(dp, 6, 2 * k + q / 2, eo, (uint32_t)S) + (q & 1)] =
                            f[4 * link_elem(dp, 3, k, eo, (uint32_t)S) + q];
    } else {
        CL2_HIP(c, hipMemcpyAsync(stage.data(), dev, stage.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
    }
    for (size_t s = 0; s < V; ++s) {
        const int par = site_parity(c, s);
        const size_t eo = s / 2;
        for (int mu = 0; mu < 4; ++mu) {
            double* u = host + 18 * (mu + 4 * s);
            for (int k = 0; k < lanes; ++k) {
                const double* d = &stage[2 * link_elem(mu * 2 + par, lanes, k, (uint32_t)eo, (uint32_t)S)];
                u[2 * k]        = d[0];
                u[2 * k + 1]    = d[1];
            }
            if (lanes == 6) {  // third row = conj(row0 x row1)
                auto cm = [](const double* a, const double* b, double* r) {
                    r[0] = a[0] * b[0] - a[1] * b[1];
                    r[1] = a[0] * b[1] + a[1] * b[0];
                };
                const int idx[3][2] = {{1, 2}, {2, 0}, {0, 1}};
                for (int col = 0; col < 3; ++col) {
                    double p1[2], p2[2];
                    cm(&u[2 * idx[col][0]], &u[2 * (3 + idx[col][1])], p1);
                    cm(&u[2 * idx[col][1]], &u[2 * (3 + idx[col][0])], p2);
                    u[2 * (6 + col)]     = p1[0] - p2[0];
                    u[2 * (6 + col) + 1] = -(p1[1] - p2[1]);
                }
            }
        }
    }
    return CL2_OK;
}

int cl2_scalar_store(cl2_ctx* c, double* dev, const double* host, int n)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dev && host && n > 0, "bad scalar store");
    CL2_HIP(c, hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    return CL2_OK;
}

int cl2_scalar_get(cl2_ctx* c, double* host, const double* dev, int n)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dev && host && n > 0, "bad scalar get");
    CL2_HIP(c, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CL2_HIP(c, hipStreamSynchronize(c->stream));
    return CL2_OK;
}

/* ---------------- Wilson ---------------- */

int cl2_dslash_eo(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt, int evenodd, double kappa)
{
    if (!c) return CL2_E_ARG;
    const double one[2] = {1., 0.};
    return run_dslash(c, out, in, nullptr, gauge, fmt, evenodd, kappa, EPI_PLAIN, one, one, 1., 0);
}

int cl2_dslash_AND_M_tm_inverse_sitediagonal_eo(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt,
                                                int evenodd, double kappa, double mubar, int minus)
{
    if (!c) return CL2_E_ARG;
    double f1[2], f2[2], s;
    tm_factors(minus ? CL2_MTM_INVERSE_SITEDIAGONAL_MINUS : CL2_MTM_INVERSE_SITEDIAGONAL, mubar, f1, f2, &s);
    return run_dslash(c, out, in, nullptr, gauge, fmt, evenodd, kappa, EPI_MINV, f1, f2, s, 0);
}

int cl2_dslash_AND_sitediagonal_saxpy_eo(cl2_ctx* c, void* out, const void* in, const void* x, const void* gauge,
                                         int fmt, int evenodd, double kappa, double f1_re, double f1_im, double f2_re,
                                         double f2_im, int gamma5)
{
    if (!c) return CL2_E_ARG;
    const double f1[2] = {f1_re, f1_im}, f2[2] = {f2_re, f2_im};
    return run_dslash(c, out, in, x, gauge, fmt, evenodd, kappa, EPI_AEE, f1, f2, 1., gamma5 ? 1 : 0);
}

int cl2_M_tm_sitediagonal_eo(cl2_ctx* c, void* out, const void* in, int prec, double mubar, int variant)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && in, "null field pointer");
    CL2_REQUIRE(c, valid_wprec(prec), "unknown precision");
    CL2_REQUIRE(c, variant >= CL2_MTM_SITEDIAGONAL && variant <= CL2_MTM_SITEDIAGONAL_MINUS_AND_GAMMA5,
                "unknown site-diagonal variant");
    CL2_PAIRS_ALIGNED(c, prec, out, in);
    double f1[2], f2[2], s;
    tm_factors(variant, mubar, f1, f2, &s);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_sitediag_h(c, out, in, f1, f2, s));
    else
        CL2_HIP(c, launch_sitediag(c, out, in, prec, f1, f2, s));
    return CL2_OK;
}

int cl2_gamma5_eo(cl2_ctx* c, void* inout, int prec)
{
    if (!c) return CL2_E_ARG;
    return run_blas(c, B_NEG_TAIL, (size_t)12 * c->g.veo, inout, inout, nullptr, nullptr, prec, value(0.), value(0.),
                    nullptr, (size_t)6 * c->g.veo);
}

int cl2_saxpy_AND_gamma5_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    CL2_REQUIRE(c, y != nullptr, "null field pointer");
    return run_blas(c, B_G5_Y_MINUS_AX, (size_t)12 * c->g.veo, out, x, y, nullptr, prec, *alpha, value(0.), nullptr,
                    (size_t)6 * c->g.veo);
}

int cl2_fermionmatrix_eo(cl2_ctx* c, int op, void* out, const void* in, const void* gauge, int fmt, int action,
                         double kappa, double mubar, void* tmp)
{
    if (!c) return CL2_E_ARG;
    return run_fermionmatrix(c, op, out, in, gauge, fmt, action, kappa, mubar, tmp, 1);
}

int cl2_fermionmatrix_eo_dot(cl2_ctx* c, int op, void* out, const void* in, const void* gauge, int fmt, int action,
                             double kappa, double mubar, void* tmp, double* dot)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, op == CL2_OP_QPLUSQMINUS_EO, "<in, A in> with the operator is for QplusQminus_eo");
    CL2_REQUIRE(c, dot != nullptr, "null result pointer");
    return run_fermionmatrix(c, op, out, in, gauge, fmt, action, kappa, mubar, tmp, 1, nullptr, nullptr, dot);
}

int cl2_fermionmatrix_eo_cg_step(cl2_ctx* c, int op, void* r, const void* p, const void* gauge, int fmt, int action,
                                 double kappa, double mubar, void* tmp, const double* r2, double* pAp, double* alpha,
                                 double* norms)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, op == CL2_OP_QPLUSQMINUS_EO, "the fused CG step is for QplusQminus_eo");
    CL2_REQUIRE(c, action == CL2_ACTION_WILSON || action == CL2_ACTION_TWISTEDMASS,
                "unknown fermion action: wilson or twistedmass");
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, r && p && gauge && tmp && r2 && pAp && alpha && norms, "null pointer");
    CL2_REQUIRE(c, r != p, "r and p must be different fields");
    const size_t fb = spinor_bytes(c, fmt_prec(fmt));
    void* t0        = tmp;
    void* t1        = static_cast<char*>(tmp) + fb;
    CL2_REQUIRE(c, r != t0 && r != t1 && p != t0 && p != t1, "tmp must not alias r or p");
    // t1 = Qminus p with pAp = |t1|^2; alpha = r2 / pAp; then Qplus t1 is folded into r <- r - alpha Qplus t1
    if (int rc = run_aee(c, t1, p, t0, gauge, fmt, action, kappa, mubar, true, true, 1, nullptr, nullptr, pAp))
        return rc;
    CL2_HIP(c, launch_scalar_op(c, CL2_SCALAR_OP_RATIO, 1, alpha, r2, pAp));
    return run_aee(c, nullptr, t1, t0, gauge, fmt, action, kappa, mubar, false, true, 1, nullptr, nullptr, norms, r,
                   alpha);
}

int cl2_fermionmatrix_eo_mrhs_cg_step(cl2_ctx* c, int op, void* r, const void* p, const void* gauge, int fmt,
                                      int action, double kappa, double mubar, const double* kappas,
                                      const double* mubars, void* tmp, int nrhs, const double* r2,
                                      const double* active, double* pAp, double* alpha, double* norms)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, op == CL2_OP_QPLUSQMINUS_EO, "the fused CG step is for QplusQminus_eo");
    CL2_REQUIRE(c, action == CL2_ACTION_WILSON || action == CL2_ACTION_TWISTEDMASS,
                "unknown fermion action: wilson or twistedmass");
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, supported_nrhs(nrhs) && nrhs > 1, "unsupported number of right-hand sides (2, 4, 8, 12 or 16)");
    CL2_REQUIRE(c, r && p && gauge && tmp && r2 && active && pAp && alpha && norms, "null pointer");
    CL2_REQUIRE(c, (kappas == nullptr) == (mubars == nullptr) || action == CL2_ACTION_WILSON,
                "kappa and mubar per right-hand side come together");
    CL2_REQUIRE(c, r != p, "r and p must be different blocks");
    const size_t fb = spinor_bytes(c, fmt_prec(fmt)) * (size_t)nrhs;
    void* t0        = tmp;
    void* t1        = static_cast<char*>(tmp) + fb;
    CL2_REQUIRE(c, r != t0 && r != t1 && p != t0 && p != t1, "tmp must not alias r or p");
    if (int rc = run_aee(c, t1, p, t0, gauge, fmt, action, kappa, mubar, true, true, nrhs, kappas, mubars, pAp))
        return rc;
    CL2_HIP(c, launch_block_cg_scalars(c, 0, alpha, const_cast<double*>(r2), pAp, active, nrhs));
    return run_aee(c, nullptr, t1, t0, gauge, fmt, action, kappa, mubar, false, true, nrhs, kappas, mubars, norms, r,
                   alpha);
}

int cl2_fermionmatrix_eo_mrhs(cl2_ctx* c, int op, void* out, const void* in, const void* gauge, int fmt, int action,
                              double kappa, double mubar, void* tmp, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, supported_nrhs(nrhs), "unsupported number of right-hand sides (1, 2, 4, 8, 12 or 16)");
    return run_fermionmatrix(c, op, out, in, gauge, fmt, action, kappa, mubar, tmp, nrhs);
}

int cl2_fermionmatrix_eo_mrhs_params(cl2_ctx* c, int op, void* out, const void* in, const void* gauge, int fmt,
                                     int action, const double* kappa, const double* mubar, void* tmp, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, supported_nrhs(nrhs) && nrhs > 1, "unsupported number of right-hand sides (2, 4, 8, 12 or 16)");
    CL2_REQUIRE(c, kappa != nullptr && (action == CL2_ACTION_WILSON || mubar != nullptr), "null parameter array");
    return run_fermionmatrix(c, op, out, in, gauge, fmt, action, 0., 0., tmp, nrhs, kappa, mubar);
}

int cl2_dslash_eo_mrhs(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt, int evenodd, double kappa,
                       int nrhs)
{
    if (!c) return CL2_E_ARG;
    const double one[2] = {1., 0.};
    return run_dslash(c, out, in, nullptr, gauge, fmt, evenodd, kappa, EPI_PLAIN, one, one, 1., 0, nrhs);
}

/* ---------------- Wilson BLAS ---------------- */

int cl2_saxpy_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    CL2_REQUIRE(c, y != nullptr, "null field pointer");
    return run_blas(c, B_Y_MINUS_AX, (size_t)12 * c->g.veo, out, x, y, nullptr, prec, *alpha, value(0.), nullptr, 0);
}

int cl2_sax_eo(cl2_ctx* c, void* out, const void* x, int prec, const cl2_scalar* alpha)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    return run_blas(c, B_AX, (size_t)12 * c->g.veo, out, x, nullptr, nullptr, prec, *alpha, value(0.), nullptr, 0);
}

int cl2_saxsbypz_eo(cl2_ctx* c, void* out, const void* x, const void* y, const void* z, int prec,
                    const cl2_scalar* alpha, const cl2_scalar* beta)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    if (int rc = check_scalar(c, beta)) return rc;
    CL2_REQUIRE(c, y && z, "null field pointer");
    return run_blas(c, B_AX_BY_Z, (size_t)12 * c->g.veo, out, x, y, z, prec, *alpha, *beta, nullptr, 0);
}

int cl2_scalar_product_eo(cl2_ctx* c, double* res, const void* x, const void* y, int prec)
{
    if (!c) return CL2_E_ARG;
    return run_reduce(c, res, x, y, (size_t)12 * c->g.veo, prec, R_DOT, 1);
}

int cl2_scalar_product_eo_noalloc(cl2_ctx* c, double* res, const void* x, const void* y, int prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_prec(prec), "cl2_scalar_product_eo_noalloc: FP64 or FP32 fields");
    CL2_REQUIRE(c, res && x && y, "null pointer");
    double* parts = scratch_ptr(c->device);
    CL2_REQUIRE(c, parts != nullptr, "no reduction scratch on this device");
    const size_t n    = (size_t)12 * c->g.veo;
    const unsigned nb = reduce_stage1_blocks(n, prec, x, y);
    CL2_REQUIRE(c, nb <= NOALLOC_PARTIALS, "cl2_scalar_product_eo_noalloc: more partial sums than the scratch holds");
    unsigned written = 0;
    CL2_HIP(c, launch_reduce_stage1(c, x, y, n, prec, R_DOT, &written, parts));
    CL2_HIP(c, launch_reduce_stage2(c, written, res, 1, nullptr, 0, parts));
    return CL2_OK;
}

int cl2_squarenorm_eo(cl2_ctx* c, double* res, const void* x, int prec)
{
    if (!c) return CL2_E_ARG;
    return run_reduce(c, res, x, nullptr, (size_t)12 * c->g.veo, prec, R_NORM, 0);
}

int cl2_saxpy_AND_squarenorm_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec,
                                const cl2_scalar* alpha, double* res)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    CL2_REQUIRE(c, y && res, "null pointer");
    return run_blas(c, B_Y_MINUS_AX, (size_t)12 * c->g.veo, out, x, y, nullptr, prec, *alpha, value(0.), res, 0);
}

int cl2_saxpy_AND_cgnorm_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha,
                            double* res)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    CL2_REQUIRE(c, y && res, "null pointer");
    return run_blas(c, B_Y_MINUS_AX, (size_t)12 * c->g.veo, out, x, y, nullptr, prec, *alpha, value(0.), res, 0, 1);
}

int cl2_cg_update_xp(cl2_ctx* c, int kind, void* x, void* p, const void* r, int prec, const cl2_scalar* alpha,
                     const cl2_scalar* beta)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    if (int rc = check_scalar(c, beta)) return rc;
    CL2_REQUIRE(c, kind == CL2_FIELD_SPINOR_EO || kind == CL2_FIELD_SU3VEC_EO, "kind must be a spinor or su3vec field");
    CL2_REQUIRE(c, x && p && r, "null field pointer");
    CL2_REQUIRE(c, x != p && x != r && p != r, "x, p and r must be different fields");
    const size_t n = (size_t)(kind == CL2_FIELD_SPINOR_EO ? 12 : 3) * c->g.veo;
    // out = z - a x on (x <- p, y <- r, z <- x), out2 = y - b x
    return run_blas(c, B_CG_XP, n, x, p, r, x, prec, *alpha, *beta, nullptr, 0, 0, p);
}

int cl2_gaussian_eo(cl2_ctx* c, int kind, void* out, int prec, uint64_t seed)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, kind == CL2_FIELD_SPINOR_EO || kind == CL2_FIELD_SU3VEC_EO, "kind must be a spinor or su3vec field");
    CL2_REQUIRE(c, out && valid_prec(prec), "bad argument");
    if (kind == CL2_FIELD_SPINOR_EO) CL2_PAIRS_ALIGNED(c, prec, out);
    CL2_HIP(c, launch_gaussian(c, out, prec, (size_t)(kind == CL2_FIELD_SPINOR_EO ? 12 : 3) * c->g.veo, seed,
                               kind == CL2_FIELD_SPINOR_EO));
    return CL2_OK;
}

int cl2_slice_squarenorm_eo(cl2_ctx* c, double* res, const void* x, int prec, int dir, int accumulate)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, res && x && valid_prec(prec), "bad argument");
    CL2_REQUIRE(c, dir == 0 || dir == 3, "correlator direction must be 0 (t) or 3 (z)");
    CL2_HIP(c, launch_slice_norm(c, res, x, prec, dir, accumulate ? 1 : 0));
    return CL2_OK;
}

int cl2_set_zero_eo(cl2_ctx* c, void* out, int prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && valid_wprec(prec), "bad argument");
    if (prec == CL2_F16)
        CL2_HIP(c, launch_fill_h(c, out, 0., 0.));
    else
        CL2_HIP(c, launch_fill(c, out, (size_t)12 * c->g.veo, prec, 0., 0.));
    return CL2_OK;
}

int cl2_set_cold_eo(cl2_ctx* c, void* out, int prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && valid_wprec(prec), "bad argument");
    const double norm = 1. / std::sqrt((double)(12. * 2. * c->g.veo));
    if (prec == CL2_F16)
        CL2_HIP(c, launch_fill_h(c, out, norm, 0.));
    else
        CL2_HIP(c, launch_fill(c, out, (size_t)12 * c->g.veo, prec, norm, 0.));
    return CL2_OK;
}

int cl2_convert_precision_eo(cl2_ctx* c, int kind, void* dst, int dst_prec, const void* src, int src_prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, kind == CL2_FIELD_SPINOR_EO || kind == CL2_FIELD_SU3VEC_EO, "kind must be a spinor or su3vec field");
    CL2_REQUIRE(c, dst && src && dst != src, "bad pointers");
    if (kind == CL2_FIELD_SPINOR_EO)
        CL2_REQUIRE(c, valid_wprec(dst_prec) && valid_wprec(src_prec), "unknown precision");
    else
        CL2_REQUIRE(c, valid_prec(dst_prec) && valid_prec(src_prec), "unknown precision (FP16 is for Wilson spinors)");
    const size_t n = (size_t)(kind == CL2_FIELD_SPINOR_EO ? 12 : 3) * c->g.veo;
    if (kind == CL2_FIELD_SPINOR_EO) {
        CL2_PAIRS_ALIGNED(c, dst_prec, dst);
        CL2_PAIRS_ALIGNED(c, src_prec, src);
    }
    if (dst_prec == src_prec)
        CL2_HIP(c, hipMemcpyAsync(dst, src, kind == CL2_FIELD_SPINOR_EO ? spinor_bytes(c, dst_prec) : su3vec_bytes(c, dst_prec),
                                  hipMemcpyDeviceToDevice, c->stream));
    else if (dst_prec == CL2_F16 || src_prec == CL2_F16)
        CL2_HIP(c, launch_convert_h(c, dst, dst_prec, src, src_prec));
    else if (kind == CL2_FIELD_SPINOR_EO)
        CL2_HIP(c, launch_convert_spinor(c, dst, dst_prec, src, src_prec, c->g.veo));
    else
        CL2_HIP(c, launch_convert(c, dst, dst_prec, src, src_prec, n));
    return CL2_OK;
}

/* ---------------- blocks of right-hand sides ---------------- */

#define CL2_BLOCK_ARGS(prec, nrhs)                                                                        \
    CL2_REQUIRE(c, valid_prec(prec), "block fields are FP64 or FP32");                                    \
    CL2_REQUIRE(c, (nrhs) >= 1 && (nrhs) <= 64, "number of right-hand sides must be in [1, 64]")
/* Wilson blocks may also be FP16 (one FP16 field of Veo nrhs sites) */
#define CL2_WBLOCK_ARGS(prec, nrhs)                                                                       \
    CL2_REQUIRE(c, valid_wprec(prec), "Wilson block fields are FP64, FP32 or FP16");                      \
    CL2_REQUIRE(c, (nrhs) >= 1 && (nrhs) <= 64, "number of right-hand sides must be in [1, 64]")

int cl2_block_insert_eo(cl2_ctx* c, void* block, const void* single, int prec, int nrhs, int r)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, block && single && r >= 0 && r < nrhs, "bad argument");
    CL2_PAIRS_ALIGNED(c, prec, block, single);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_block_slot_h(c, block, single, nullptr, nrhs, r));
    else
        CL2_HIP(c, launch_block_slot(c, block, single, nullptr, prec, nrhs, r));
    return CL2_OK;
}

int cl2_block_extract_eo(cl2_ctx* c, void* single, const void* block, int prec, int nrhs, int r)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, block && single && r >= 0 && r < nrhs, "bad argument");
    CL2_PAIRS_ALIGNED(c, prec, block, single);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_block_slot_h(c, const_cast<void*>(block), nullptr, single, nrhs, r));
    else
        CL2_HIP(c, launch_block_slot(c, const_cast<void*>(block), nullptr, single, prec, nrhs, r));
    return CL2_OK;
}

int cl2_block_axpy_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int nrhs, const double* coef,
                      double sign, double* norms)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, out && x && y && coef, "null pointer");
    CL2_PAIRS_ALIGNED(c, prec, out, x, y);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_block_axpy_h(c, out, x, y, nrhs, coef, sign, norms));
    else
        CL2_HIP(c, launch_block_axpy(c, out, x, y, prec, nrhs, coef, sign, norms));
    return CL2_OK;
}

int cl2_block_cg_update_xp_eo(cl2_ctx* c, void* x, void* p, const void* r, int prec, int nrhs, const double* alpha,
                              const double* beta)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, x && p && r && alpha && beta, "null pointer");
    CL2_REQUIRE(c, x != p && x != r && p != r, "x, p and r must be different blocks");
    CL2_PAIRS_ALIGNED(c, prec, x, p, r);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_block_xp_h(c, x, p, r, nrhs, alpha, beta));
    else
        CL2_HIP(c, launch_block_xp(c, x, p, r, prec, nrhs, alpha, beta));
    return CL2_OK;
}

int cl2_block_dot_eo(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, res && x && y, "null pointer");
    CL2_PAIRS_ALIGNED(c, prec, x, y);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_block_dot_h(c, res, x, y, nrhs, 0));
    else
        CL2_HIP(c, launch_block_dot(c, res, x, y, prec, nrhs, 0));
    return CL2_OK;
}

int cl2_block_squarenorm_eo(cl2_ctx* c, double* res, const void* x, int prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, res && x, "null pointer");
    CL2_PAIRS_ALIGNED(c, prec, x);
    if (prec == CL2_F16)
        CL2_HIP(c, launch_block_dot_h(c, res, x, x, nrhs, 1));
    else
        CL2_HIP(c, launch_block_dot(c, res, x, x, prec, nrhs, 1));
    return CL2_OK;
}

int cl2_block_convert_eo(cl2_ctx* c, void* dst, int dst_prec, const void* src, int src_prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(dst_prec, nrhs);
    CL2_REQUIRE(c, valid_wprec(src_prec), "Wilson block fields are FP64, FP32 or FP16");
    CL2_REQUIRE(c, dst && src && dst != src, "bad pointers");
    CL2_PAIRS_ALIGNED(c, dst_prec, dst);
    CL2_PAIRS_ALIGNED(c, src_prec, src);
    const uint32_t sites = c->g.veo * (uint32_t)nrhs;
    if (dst_prec == src_prec)
        CL2_HIP(c, hipMemcpyAsync(dst, src, spinor_bytes(c, dst_prec) * (size_t)nrhs, hipMemcpyDeviceToDevice,
                                  c->stream));
    else if (dst_prec == CL2_F16 || src_prec == CL2_F16)
        CL2_HIP(c, launch_convert_h(c, dst, dst_prec, src, src_prec, sites));
    else
        CL2_HIP(c, launch_convert_spinor(c, dst, dst_prec, src, src_prec, (size_t)sites));
    return CL2_OK;
}

int cl2_block_set_zero_eo(cl2_ctx* c, void* out, int prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_WBLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, out != nullptr, "null pointer");
    if (prec == CL2_F16)
        CL2_HIP(c, launch_fill_h(c, out, 0., 0., c->g.veo * (uint32_t)nrhs));
    else
        CL2_HIP(c, launch_fill(c, out, (size_t)12 * c->g.veo * nrhs, prec, 0., 0.));
    return CL2_OK;
}

/* ---- blocks of staggered vectors [3][Veo][nrhs] ---- */
int cl2_block_insert_stagg_eo(cl2_ctx* c, void* block, const void* single, int prec, int nrhs, int r)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, block && single && r >= 0 && r < nrhs, "bad argument");
    CL2_HIP(c, launch_block_slot(c, block, single, nullptr, prec, nrhs, r, 3));
    return CL2_OK;
}

int cl2_block_extract_stagg_eo(cl2_ctx* c, void* single, const void* block, int prec, int nrhs, int r)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, block && single && r >= 0 && r < nrhs, "bad argument");
    CL2_HIP(c, launch_block_slot(c, const_cast<void*>(block), nullptr, single, prec, nrhs, r, 3));
    return CL2_OK;
}

int cl2_block_axpy_stagg_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, int nrhs,
                            const double* coef, double sign, double* norms)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, out && x && y && coef, "null pointer");
    CL2_HIP(c, launch_block_axpy(c, out, x, y, prec, nrhs, coef, sign, norms, 3));
    return CL2_OK;
}

int cl2_block_dot_stagg_eo(cl2_ctx* c, double* res, const void* x, const void* y, int prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, res && x && y, "null pointer");
    CL2_HIP(c, launch_block_dot(c, res, x, y, prec, nrhs, 0, 3));
    return CL2_OK;
}

int cl2_block_squarenorm_stagg_eo(cl2_ctx* c, double* res, const void* x, int prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, res && x, "null pointer");
    CL2_HIP(c, launch_block_dot(c, res, x, x, prec, nrhs, 1, 3));
    return CL2_OK;
}

int cl2_block_convert_stagg_eo(cl2_ctx* c, void* dst, int dst_prec, const void* src, int src_prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(dst_prec, nrhs);
    CL2_REQUIRE(c, valid_prec(src_prec), "block fields are FP64 or FP32");
    CL2_REQUIRE(c, dst && src && dst != src, "bad pointers");
    const size_t n = (size_t)3 * c->g.veo * nrhs;
    if (dst_prec == src_prec)
        CL2_HIP(c, hipMemcpyAsync(dst, src, n * (dst_prec == CL2_F64 ? 16 : 8), hipMemcpyDeviceToDevice, c->stream));
    else
        CL2_HIP(c, launch_convert(c, dst, dst_prec, src, src_prec, n));
    return CL2_OK;
}

int cl2_block_set_zero_stagg_eo(cl2_ctx* c, void* out, int prec, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, out != nullptr, "null pointer");
    CL2_HIP(c, launch_fill(c, out, (size_t)3 * c->g.veo * nrhs, prec, 0., 0.));
    return CL2_OK;
}

int cl2_block_cg_update_xp_stagg_eo(cl2_ctx* c, void* x, void* p, const void* r, int prec, int nrhs,
                                    const double* alpha, const double* beta)
{
    if (!c) return CL2_E_ARG;
    CL2_BLOCK_ARGS(prec, nrhs);
    CL2_REQUIRE(c, x && p && r && alpha && beta, "null pointer");
    CL2_REQUIRE(c, x != p && x != r && p != r, "x, p and r must be different blocks");
    CL2_HIP(c, launch_block_xp(c, x, p, r, prec, nrhs, alpha, beta, 3));
    return CL2_OK;
}

int cl2_D_KS_eo_mrhs(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt, int evenodd, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, supported_nrhs(nrhs) && nrhs > 1, "unsupported number of right-hand sides (2, 4, 8, 12 or 16)");
    CL2_REQUIRE(c, evenodd == CL2_EVEN || evenodd == CL2_ODD, "evenodd must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, out && in && gauge, "null field pointer");
    CL2_REQUIRE(c, out != in, "D_KS_eo cannot run in place");
    CL2_HIP(c, launch_dks_mrhs(c, out, in, gauge, fmt, evenodd, nrhs, nullptr, nullptr));
    return CL2_OK;
}

int cl2_MdagM_eo_mrhs(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt, double mass, int upper_left,
                      const double* sigma, void* tmp, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, supported_nrhs(nrhs) && nrhs > 1, "unsupported number of right-hand sides (2, 4, 8, 12 or 16)");
    CL2_REQUIRE(c, upper_left == CL2_EVEN || upper_left == CL2_ODD, "upper_left must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, out && in && gauge && tmp, "null field pointer");
    CL2_REQUIRE(c, out != in && tmp != in && tmp != out, "out, in and tmp must be different blocks");
    double coef[16];
    for (int r = 0; r < nrhs; ++r) coef[r] = mass * mass + (sigma ? sigma[r] : 0.);
    const int first = (upper_left == CL2_EVEN) ? CL2_ODD : CL2_EVEN;
    CL2_HIP(c, launch_dks_mrhs(c, tmp, in, gauge, fmt, first, nrhs, nullptr, nullptr));
    CL2_HIP(c, launch_dks_mrhs(c, out, tmp, gauge, fmt, upper_left, nrhs, in, coef));
    return CL2_OK;
}

int cl2_MdagM_eo_mrhs_cg_step(cl2_ctx* c, void* r, const void* p, const void* gauge, int fmt, double mass,
                              int upper_left, const double* sigma, void* tmp, int nrhs, const double* r2,
                              const double* active, double* pAp, double* alpha, double* norms)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, supported_nrhs(nrhs) && nrhs > 1, "unsupported number of right-hand sides (2, 4, 8, 12 or 16)");
    CL2_REQUIRE(c, upper_left == CL2_EVEN || upper_left == CL2_ODD, "upper_left must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, r && p && gauge && tmp && r2 && active && pAp && alpha && norms, "null pointer");
    CL2_REQUIRE(c, r != p && tmp != p && tmp != r, "r, p and tmp must be different blocks");
    double coef[16];
    for (int k = 0; k < nrhs; ++k) coef[k] = mass * mass + (sigma ? sigma[k] : 0.);
    const int first = (upper_left == CL2_EVEN) ? CL2_ODD : CL2_EVEN;
    double2* parts  = reinterpret_cast<double2*>(c->partials);
    unsigned nb     = 0;
    // w = D_KS p with per-right-hand-side (|w|^2, |p|^2); pAp and alpha on the device; then r -= alpha (coef p - D w)
    CL2_HIP(c, launch_dks_mrhs(c, tmp, p, gauge, fmt, first, nrhs, nullptr, nullptr, parts, p, nullptr, nullptr, &nb));
    CL2_HIP(c, launch_block_reduce2(c, parts, nb, pAp, 0, nrhs));
    CL2_HIP(c, launch_block_cg_alpha_mdagm(c, alpha, r2, pAp, coef, active, nrhs));
    CL2_HIP(c, launch_dks_mrhs(c, nullptr, tmp, gauge, fmt, upper_left, nrhs, p, coef, parts, nullptr, r, alpha, &nb));
    CL2_HIP(c, launch_block_reduce2(c, parts, nb, norms, 0, nrhs));
    return CL2_OK;
}

int cl2_block_cg_alpha(cl2_ctx* c, double* alpha, const double* r2, const double* pAp, const double* active,
                       int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, alpha && r2 && pAp && active && nrhs >= 1 && nrhs <= 64, "bad argument");
    CL2_HIP(c, launch_block_cg_scalars(c, 0, alpha, const_cast<double*>(r2), pAp, active, nrhs));
    return CL2_OK;
}

int cl2_block_cg_beta(cl2_ctx* c, double* beta, double* r2, const double* norms, const double* active, int nrhs)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, beta && r2 && norms && active && nrhs >= 1 && nrhs <= 64, "bad argument");
    CL2_HIP(c, launch_block_cg_scalars(c, 1, beta, r2, norms, active, nrhs));
    return CL2_OK;
}

/* ---------------- multi-shift ---------------- */

int cl2_multishift_update_stagg_eo(cl2_ctx* c, void* const* x, int x_prec, void* const* p, const void* r, int prec,
                                   int nshift, const double* beta, const double* zeta, const double* alpha,
                                   const double* active)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_prec(prec) && valid_prec(x_prec), "unknown precision");
    CL2_REQUIRE(c, x_prec == CL2_F64 || prec == CL2_F32, "x must be at least as precise as r and p");
    CL2_REQUIRE(c, nshift >= 1 && nshift <= 64, "number of shifts must be in [1, 64]");
    CL2_REQUIRE(c, x && p && r && beta && zeta && alpha && active, "null pointer");
    std::vector<void*> tab(128, nullptr);
    for (int k = 0; k < nshift; ++k) {
        CL2_REQUIRE(c, x[k] && p[k] && x[k] != p[k], "null or aliased shift field");
        tab[k]      = x[k];
        tab[64 + k] = p[k];
    }
    if (tab != c->ptr_table_host) {  // upload only when the fields change
        CL2_HIP(c, hipMemcpyAsync(c->ptr_table, tab.data(), 128 * sizeof(void*), hipMemcpyHostToDevice, c->stream));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
        c->ptr_table_host = tab;
    }
    CL2_HIP(c, launch_multishift(c, c->ptr_table, x_prec, c->ptr_table + 64, r, prec, (size_t)3 * c->g.veo, nshift, beta,
                                 zeta, alpha, active));
    return CL2_OK;
}

/* ---------------- staggered ---------------- */

int cl2_D_KS_eo(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt, int evenodd)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_fmt(fmt) && fmt != CL2_GAUGE_F16_12, "unknown gauge format (F16_12 is for Wilson operators)");
    CL2_REQUIRE(c, evenodd == CL2_EVEN || evenodd == CL2_ODD, "evenodd must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, out && in && gauge, "null field pointer");
    CL2_REQUIRE(c, out != in, "D_KS_eo cannot run in place: out aliases in");
    CL2_HIP(c, launch_dks(c, out, in, gauge, fmt, evenodd));
    return CL2_OK;
}

int cl2_MdagM_eo(cl2_ctx* c, void* out, const void* in, const void* gauge, int fmt, double mass, int upper_left,
                 void* tmp)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, tmp && tmp != in && tmp != out, "tmp must be a separate field");
    CL2_REQUIRE(c, upper_left == CL2_EVEN || upper_left == CL2_ODD, "upper_left must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, out != in, "MdagM_eo cannot run in place");
    const int first = (upper_left == CL2_EVEN) ? CL2_ODD : CL2_EVEN;
    int rc          = cl2_D_KS_eo(c, tmp, in, gauge, fmt, first);
    if (rc) return rc;
    rc = cl2_D_KS_eo(c, out, tmp, gauge, fmt, upper_left);
    if (rc) return rc;
    return run_blas(c, B_AX_BY, (size_t)3 * c->g.veo, out, in, out, nullptr, fmt_prec(fmt), value(mass * mass),
                    value(-1.), nullptr, 0);
}

int cl2_saxpy_stagg_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    CL2_REQUIRE(c, y != nullptr, "null field pointer");
    return run_blas(c, B_AX_PLUS_Y, (size_t)3 * c->g.veo, out, x, y, nullptr, prec, *alpha, value(0.), nullptr, 0);
}

int cl2_saxpby_stagg_eo(cl2_ctx* c, void* out, const void* x, const void* y, int prec, const cl2_scalar* alpha,
                        const cl2_scalar* beta)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    if (int rc = check_scalar(c, beta)) return rc;
    CL2_REQUIRE(c, y != nullptr, "null field pointer");
    return run_blas(c, B_AX_BY, (size_t)3 * c->g.veo, out, x, y, nullptr, prec, *alpha, *beta, nullptr, 0);
}

int cl2_sax_stagg_eo(cl2_ctx* c, void* out, const void* x, int prec, const cl2_scalar* alpha)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    return run_blas(c, B_AX, (size_t)3 * c->g.veo, out, x, nullptr, nullptr, prec, *alpha, value(0.), nullptr, 0);
}

int cl2_saxpbypz_stagg_eo(cl2_ctx* c, void* out, const void* x, const void* y, const void* z, int prec,
                          const cl2_scalar* alpha, const cl2_scalar* beta)
{
    if (!c) return CL2_E_ARG;
    if (int rc = check_scalar(c, alpha)) return rc;
    if (int rc = check_scalar(c, beta)) return rc;
    CL2_REQUIRE(c, y && z, "null field pointer");
    return run_blas(c, B_AX_BY_Z, (size_t)3 * c->g.veo, out, x, y, z, prec, *alpha, *beta, nullptr, 0);
}

int cl2_scalar_product_stagg_eo(cl2_ctx* c, double* res, const void* x, const void* y, int prec)
{
    if (!c) return CL2_E_ARG;
    return run_reduce(c, res, x, y, (size_t)3 * c->g.veo, prec, R_DOT, 1);
}

int cl2_scalar_product_real_part_stagg_eo(cl2_ctx* c, double* res, const void* x, const void* y, int prec)
{
    if (!c) return CL2_E_ARG;
    return run_reduce(c, res, x, y, (size_t)3 * c->g.veo, prec, R_DOT_RE, 0);
}

int cl2_squarenorm_stagg_eo(cl2_ctx* c, double* res, const void* x, int prec)
{
    if (!c) return CL2_E_ARG;
    return run_reduce(c, res, x, nullptr, (size_t)3 * c->g.veo, prec, R_NORM, 0);
}

int cl2_sax_vec_AND_squarenorm_stagg_eo(cl2_ctx* c, double* res, const double* alpha, int n, const void* x, int prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, res && alpha && x && n > 0, "bad argument");
    CL2_REQUIRE(c, valid_prec(prec), "unknown precision");
    unsigned nb = 0;
    CL2_HIP(c, launch_reduce_stage1(c, x, nullptr, (size_t)3 * c->g.veo, prec, R_NORM, &nb));
    CL2_HIP(c, launch_reduce_stage2(c, nb, res, 2, alpha, n));
    return CL2_OK;
}

int cl2_set_zero_stagg_eo(cl2_ctx* c, void* out, int prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && valid_prec(prec), "bad argument");
    CL2_HIP(c, launch_fill(c, out, (size_t)3 * c->g.veo, prec, 0., 0.));
    return CL2_OK;
}

int cl2_set_cold_stagg_eo(cl2_ctx* c, void* out, int prec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && valid_prec(prec), "bad argument");
    const double norm = 1. / std::sqrt((double)(3. * 2. * c->g.veo));
    CL2_HIP(c, launch_fill(c, out, (size_t)3 * c->g.veo, prec, norm, 0.));
    return CL2_OK;
}

/* ---------------- device scalars ---------------- */

int cl2_complex_op(cl2_ctx* c, int op, double* out, const double* a, const double* b)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && a && b, "null scalar");
    CL2_REQUIRE(c, op >= CL2_SCALAR_OP_PRODUCT && op <= CL2_SCALAR_OP_SUBTRACTION, "unknown scalar op");
    CL2_HIP(c, launch_scalar_op(c, op, 1, out, a, b));
    return CL2_OK;
}

int cl2_real_op(cl2_ctx* c, int op, double* out, const double* a, const double* b)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && a && b, "null scalar");
    CL2_REQUIRE(c, op >= CL2_SCALAR_OP_PRODUCT && op <= CL2_SCALAR_OP_SUBTRACTION, "unknown scalar op");
    CL2_HIP(c, launch_scalar_op(c, op, 0, out, a, b));
    return CL2_OK;
}

int cl2_real_get_vector_element(cl2_ctx* c, double* out, const double* vec, int index)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && vec && index >= 0, "bad argument");
    CL2_HIP(c, launch_vec_element(c, out, vec, 0, index));
    return CL2_OK;
}

int cl2_real_set_vector_element(cl2_ctx* c, double* vec, const double* in, int index)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, in && vec && index >= 0, "bad argument");
    CL2_HIP(c, launch_vec_element(c, vec, in, index, 0));
    return CL2_OK;
}

int cl2_update_zeta_cgm(cl2_ctx* c, double* out, const double* zeta_prev, const double* zeta_prev_prev,
                        const double* sbeta_prev, const double* sbeta_pres, const double* salpha_prev,
                        const double* sigma, int numeq)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && zeta_prev && zeta_prev_prev && sbeta_prev && sbeta_pres && salpha_prev && sigma && numeq > 0,
                "bad argument");
    CL2_HIP(c, launch_cgm_update(c, 0, out, zeta_prev, zeta_prev_prev, sbeta_prev, sbeta_pres, salpha_prev, sigma,
                                 numeq));
    return CL2_OK;
}

int cl2_update_beta_cgm(cl2_ctx* c, double* out, const double* sbeta_pres, const double* zeta_pres,
                        const double* zeta_prev, int numeq)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && sbeta_pres && zeta_pres && zeta_prev && numeq > 0, "bad argument");
    CL2_HIP(c, launch_cgm_update(c, 1, out, sbeta_pres, zeta_pres, zeta_prev, nullptr, nullptr, nullptr, numeq));
    return CL2_OK;
}

int cl2_update_alpha_cgm(cl2_ctx* c, double* out, const double* salpha_pres, const double* zeta_pres,
                         const double* beta_pres, const double* zeta_prev, const double* sbeta_pres, int numeq)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out && salpha_pres && zeta_pres && beta_pres && zeta_prev && sbeta_pres && numeq > 0,
                "bad argument");
    CL2_HIP(c, launch_cgm_update(c, 2, out, salpha_pres, zeta_pres, beta_pres, zeta_prev, sbeta_pres, nullptr,
                                 numeq));
    return CL2_OK;
}

int cl2_cgm_beta_step(cl2_ctx* c, double* beta, double* beta_prev, const double* rr, const double* pap)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, beta && beta_prev && rr && pap, "bad argument");
    CL2_HIP(c, launch_cgm_beta_step(c, beta, beta_prev, rr, pap));
    return CL2_OK;
}

int cl2_MdagM_eo_cgm_step(cl2_ctx* c, void* r, const void* p, const void* gauge, int fmt, double mass, double sigma,
                          int upper_left, void* tmp, const double* rr_old, double* pap, double* beta,
                          double* beta_prev, double* rr_new)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, valid_fmt(fmt), "unknown gauge format");
    CL2_REQUIRE(c, upper_left == CL2_EVEN || upper_left == CL2_ODD, "upper_left must be CL2_EVEN or CL2_ODD");
    CL2_REQUIRE(c, r && p && gauge && tmp && rr_old && pap && beta && beta_prev && rr_new, "null pointer");
    CL2_REQUIRE(c, r != p && tmp != p && tmp != r, "r, p and tmp must be different fields");
    const int first = (upper_left == CL2_EVEN) ? CL2_ODD : CL2_EVEN;
    const double m2 = mass * mass + sigma;
    unsigned nb     = 0;
    double* wp      = c->dscratch + 2;  // (|D p|^2, |p|^2)
    CL2_HIP(c, launch_dks_step(c, tmp, p, gauge, fmt, first, p, nullptr, 0., nullptr, nullptr, &nb));
    CL2_HIP(c, launch_reduce_stage2(c, nb, wp, 1, nullptr, 0));
    CL2_HIP(c, launch_cgm_beta_step_mdagm(c, beta, beta_prev, rr_old, pap, wp, m2));
    CL2_HIP(c, launch_dks_step(c, nullptr, tmp, gauge, fmt, upper_left, nullptr, p, m2, r, beta, &nb));
    CL2_HIP(c, launch_reduce_stage2(c, nb, rr_new, 0, nullptr, 0));
    return CL2_OK;
}

int cl2_cgm_alpha_step(cl2_ctx* c, int numeq, double* alpha, double* alpha_prev, const double* rr_new,
                       double* rr_old, const double* beta, const double* beta_prev, const double* sigma, double* zeta,
                       double* zeta_prev, double* zeta_foll, double* beta_vec, double* alpha_vec)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, numeq > 0 && numeq <= 64, "cl2_cgm_alpha_step: 1 to 64 shifts");
    CL2_REQUIRE(c, alpha && alpha_prev && rr_new && rr_old && beta && beta_prev && sigma && zeta && zeta_prev &&
                       zeta_foll && beta_vec && alpha_vec,
                "bad argument");
    CL2_HIP(c, launch_cgm_alpha_step(c, numeq, alpha, alpha_prev, rr_new, rr_old, beta, beta_prev, sigma, zeta,
                                     zeta_prev, zeta_foll, beta_vec, alpha_vec));
    return CL2_OK;
}

/* ---------------- captured launch sequences ---------------- */

}  // extern "C"

struct cl2_graph {
    hipGraph_t graph    = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int cl2_graph_begin(cl2_ctx* c)
{
    if (!c) return CL2_E_ARG;
    CL2_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    return CL2_OK;
}

int cl2_graph_end(cl2_ctx* c, cl2_graph** out)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, out != nullptr, "null output pointer");
    *out = nullptr;
    hipGraph_t g = nullptr;
    CL2_HIP(c, hipStreamEndCapture(c->stream, &g));
    cl2_graph* h = new cl2_graph;
    h->graph     = g;
    if (hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete h;
        return fail(c, CL2_E_HIP, "hipGraphInstantiate failed");
    }
    *out = h;
    return CL2_OK;
}

int cl2_graph_launch(cl2_ctx* c, cl2_graph* g, int times)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, g != nullptr && times >= 0, "bad argument");
    for (int k = 0; k < times; ++k) CL2_HIP(c, hipGraphLaunch(g->exec, c->stream));
    return CL2_OK;
}

int cl2_graph_node_count(const cl2_graph* g, int* kernel_nodes, int* all_nodes)
{
    if (!g || !g->graph || !kernel_nodes || !all_nodes) return CL2_E_ARG;
    size_t n = 0;
    if (hipGraphGetNodes(g->graph, nullptr, &n) != hipSuccess) return CL2_E_HIP;
    std::vector<hipGraphNode_t> nodes(n);
    if (n && hipGraphGetNodes(g->graph, nodes.data(), &n) != hipSuccess) return CL2_E_HIP;
    int k = 0;
    for (size_t i = 0; i < n; ++i) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[i], &t) != hipSuccess) return CL2_E_HIP;
        k += t == hipGraphNodeTypeKernel;
    }
    *kernel_nodes = k;
    *all_nodes    = (int)n;
    return CL2_OK;
}

int cl2_graph_destroy(cl2_graph* g)
{
    if (!g) return CL2_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return CL2_OK;
}

int cl2_scalar_copy(cl2_ctx* c, double* dst, const double* src, int n)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, dst && src && n > 0 && n <= (1 << 20), "bad scalar copy");
    if (dst == src) return CL2_OK;
    CL2_HIP(c, launch_scalar_copy(c, dst, src, n));
    return CL2_OK;
}

int cl2_cg_beta(cl2_ctx* c, double* mbeta, double* r2, const double* norms)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, mbeta && r2 && norms, "null pointer");
    CL2_HIP(c, launch_cg_beta(c, mbeta, r2, norms));
    return CL2_OK;
}

/* ---------------- tuning ---------------- */

int cl2_tune_set(cl2_ctx* c, const char* key, long value)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, key != nullptr, "null key");
    const std::string k(key);
    if (k == "hop.zplanes") {
        CL2_REQUIRE(c, value == -1 || value == 0 || (value > 0 && c->g.ns % (uint32_t)value == 0),
                    "hop.zplanes must be -1, 0 or a divisor of ns");
        c->tune_zplanes = (int)value;
        c->tune_manual  = true;
        return CL2_OK;
    }
    if (k == "hop.xcd") {
        c->tune_xcd    = value ? 1 : 0;
        c->tune_manual = true;
        return CL2_OK;
    }
    if (k == "mrhs.rb") {
        CL2_REQUIRE(c, value == 0 || value == 1 || value == 2 || value == 4, "mrhs.rb must be 0, 1, 2 or 4");
        c->tune_mrhs_rb = (int)value;
        c->tune_manual  = true;
        return CL2_OK;
    }
    if (k == "mrhs.lds") {
        CL2_REQUIRE(c, value == 0 || value == 1 || value == 2, "mrhs.lds must be 0, 1 or 2");
        c->tune_mrhs_lds = (int)value;
        c->tune_manual   = true;
        return CL2_OK;
    }
    if (k == "mrhs.off32") {
        c->tune_mrhs_off32 = value ? 1 : 0;
        return CL2_OK;
    }
    if (k == "hop.block") {
        CL2_REQUIRE(c, value == 128 || value == 256 || value == 512, "hop.block must be 128, 256 or 512");
        c->tune_block  = (int)value;
        c->tune_manual = true;
        return CL2_OK;
    }
    if (k == "hop.occupancy") {
        CL2_REQUIRE(c, value == -1 || value == 0 || value == 1 || value == 2 || value == 4,
                    "hop.occupancy must be -1, 0, 1, 2 or 4");
        c->tune_occ    = (int)value;
        c->tune_manual = true;
        return CL2_OK;
    }
    return fail(c, CL2_E_ARG, "unknown tuning key " + k);
}

long cl2_tune_get(const cl2_ctx* c, const char* key)
{
    if (!c || !key) return -2;
    const std::string k(key);
    if (k == "hop.zplanes") return c->tune_zplanes;
    if (k == "hop.xcd") return c->tune_xcd;
    if (k == "mrhs.rb") return c->tune_mrhs_rb;
    if (k == "hop.block") return c->tune_block;
    if (k == "mrhs.lds") return c->tune_mrhs_lds;
    if (k == "mrhs.off32") return c->tune_mrhs_off32;
    if (k == "hop.occupancy") return c->tune_occ;
    return -2;
}

/* Autotuner of the hopping kernel's launch shape: every candidate (z-planes per slab, XCD mapping,
 * right-hand sides per thread) is timed with HIP events on scratch fields and the caller's gauge field;
 * the fastest is stored for (format, nrhs) and used by every later launch on this context unless
 * cl2_tune_set has fixed the shape by hand. */
int cl2_autotune_dslash(cl2_ctx* c, const void* gauge, int fmt, int nrhs, double* best_ms, int* zplanes, int* xcd,
                        int* rb, int* block, int* occupancy)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, gauge && valid_fmt(fmt), "bad gauge argument");
    CL2_REQUIRE(c, supported_nrhs(nrhs) && (nrhs == 1 || fmt != CL2_GAUGE_F16_12), "unsupported right-hand sides");
    const int prec    = fmt_prec(fmt);
    const size_t fb   = spinor_bytes(c, prec) * (size_t)nrhs;
    TempDev in, out;
    CL2_HIP(c, hipMalloc(&in.p, fb));
    CL2_HIP(c, hipMalloc(&out.p, fb));
    // Gaussian spinors (FP16: rounded from FP32 ones), data like a solver's fields
    if (prec == CL2_F16) {
        TempDev g32;
        CL2_HIP(c, hipMalloc(&g32.p, spinor_bytes(c, CL2_F32)));
        CL2_HIP(c, launch_gaussian(c, g32.p, CL2_F32, (size_t)12 * c->g.veo, 0x5eed, true));
        CL2_HIP(c, launch_convert_h(c, in.p, CL2_F16, g32.p, CL2_F32));
        CL2_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        CL2_HIP(c, launch_gaussian(c, in.p, prec, (size_t)12 * c->g.veo * nrhs, 0x5eed, nrhs == 1));
    }
    std::vector<int> zps = {-1, 0};
    for (uint32_t d = 1; d <= c->g.ns; ++d)
        if (c->g.ns % d == 0 && (size_t)d * c->g.ns * c->g.ns / 2 * nrhs <= 262144) zps.push_back((int)d);
    std::vector<int> rbs = {1};
    if (nrhs > 1)
        for (int r : {2, 4})
            if (nrhs % r == 0) rbs.push_back(r);
    // (block, occupancy hint) shapes of the single-field kernel
    std::vector<std::pair<int, int>> shapes = {{256, 0}};
    if (nrhs == 1 && fmt == CL2_GAUGE_F32_12)
        shapes = {{256, 0}, {128, 0}, {512, 0}, {256, 1}, {128, 1}, {256, 2}, {128, 2}, {256, 4}, {128, 4}};
    else if (nrhs == 1 && fmt == CL2_GAUGE_F16_12)
        shapes = {{256, 0}, {128, 0}, {512, 0}, {256, 1}, {128, 1}, {256, 4}, {128, 4}};
    else if (nrhs == 1) shapes = {{256, 0}, {128, 0}, {512, 0}, {256, 1}, {128, 1}};
    const bool manual = c->tune_manual;
    c->tune_manual    = false;
    const double one[2] = {1., 0.};
    double best         = 1e30;
    cl2_hop_tuning bt{-1, 1, 1, 1, DSLASH_BLOCK, -1};
    int rc = CL2_OK;
    // multi-RHS with one right-hand side per thread: with and without the per-site LDS staging
    std::vector<std::pair<int, int>> rls;
    for (int r : rbs) {
        rls.push_back({r, 0});
        if (nrhs > 1 && r == 1) {
            rls.push_back({r, 1});
            rls.push_back({r, 2});
        }
    }
    // ms per application of shape t over n timed launches (after two untimed ones)
    auto time_shape = [&](const cl2_hop_tuning& t, int n, double& ms_out) -> int {
        c->tuned[{fmt, nrhs}] = t;
        int e                 = CL2_OK;
        for (int k = 0; k < 2 && e == CL2_OK; ++k)
            e = run_dslash(c, out.p, in.p, nullptr, gauge, fmt, CL2_EVEN, 0.125, EPI_PLAIN, one, one, 1., 0, nrhs);
        if (e) return e;
        if (hipEventRecord(c->ev_start, c->stream) != hipSuccess) return CL2_E_HIP;
        for (int k = 0; k < n && e == CL2_OK; ++k)
            e = run_dslash(c, out.p, in.p, nullptr, gauge, fmt, k & 1 ? CL2_ODD : CL2_EVEN, 0.125, EPI_PLAIN, one,
                           one, 1., 0, nrhs);
        if (e) return e;
        float ms = 0.f;
        if (hipEventRecord(c->ev_stop, c->stream) != hipSuccess || hipEventSynchronize(c->ev_stop) != hipSuccess ||
            hipEventElapsedTime(&ms, c->ev_start, c->ev_stop) != hipSuccess)
            return CL2_E_HIP;
        ms_out = ms / n;
        return CL2_OK;
    };
    // every shape briefly; then the three fastest again over four times the launches, so that one short
    // measurement caught in a clock or power transient does not decide
    std::vector<std::pair<double, cl2_hop_tuning>> timed;
    for (int zp : zps)
        for (int x : {1, 0})
            for (const auto& rl : rls)
                for (const auto& sh : shapes) {
                    if (rc != CL2_OK) continue;  // a failed launch ends the search
                    const cl2_hop_tuning t{zp, x, rl.first, rl.second, sh.first, sh.second};
                    double ms = 0.;
                    rc        = time_shape(t, 5, ms);
                    if (rc == CL2_OK) timed.push_back({ms, t});
                }
    if (rc == CL2_OK) {
        std::stable_sort(timed.begin(), timed.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < timed.size() && k < 3 && rc == CL2_OK; ++k) {
            double ms = 0.;
            if ((rc = time_shape(timed[k].second, 20, ms)) != CL2_OK) break;
            if (ms < best) {
                best = ms;
                bt   = timed[k].second;
            }
        }
    }
    c->tune_manual = manual;
    if (rc) {
        c->tuned.erase({fmt, nrhs});
        return rc;
    }
    c->tuned[{fmt, nrhs}] = bt;
    if (best_ms) *best_ms = best;
    if (zplanes) *zplanes = bt.zplanes;
    if (xcd) *xcd = bt.xcd;
    if (rb) *rb = (nrhs > 1 && bt.rb <= 1 && bt.lds) ? -bt.lds : bt.rb;
    if (block) *block = bt.block;
    if (occupancy) *occupancy = bt.occ;
    return CL2_OK;
}

/* The same for the staggered hopping kernel: z-planes per slab x XCD mapping x block size x register
 * budget, timed on scratch staggered fields with the caller's gauge field. */
int cl2_autotune_dks(cl2_ctx* c, const void* gauge, int fmt, double* best_ms, int* zplanes, int* xcd, int* block,
                     int* occupancy)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, gauge && valid_fmt(fmt) && fmt != CL2_GAUGE_F16_12, "bad gauge argument");
    const int prec  = fmt_prec(fmt);
    const size_t fb = su3vec_bytes(c, prec);
    TempDev in, out;
    CL2_HIP(c, hipMalloc(&in.p, fb));
    CL2_HIP(c, hipMalloc(&out.p, fb));
    CL2_HIP(c, launch_fill(c, in.p, (size_t)3 * c->g.veo, prec, 0.1, 0.));
    std::vector<int> zps = {-1, 0};
    for (uint32_t d = 1; d <= c->g.ns; ++d)
        if (c->g.ns % d == 0 && (size_t)d * c->g.ns * c->g.ns / 2 <= 262144) zps.push_back((int)d);
    const std::vector<std::pair<int, int>> shapes = {{256, 0}, {128, 0}, {512, 0}, {256, 1}, {128, 1}, {256, 4},
                                                     {128, 4}};
    const std::pair<int, int> key{fmt | DKS_TUNING_KEY, 1};
    const bool manual = c->tune_manual;
    c->tune_manual    = false;
    double best       = 1e30;
    cl2_hop_tuning bt{-1, 1, 1, 1, DSLASH_BLOCK, 0};
    hipError_t e = hipSuccess;
    for (int zp : zps)
        for (int x : {1, 0})
            for (const auto& sh : shapes) {
                c->tuned[key] = cl2_hop_tuning{zp, x, 1, 1, sh.first, sh.second};
                for (int k = 0; k < 2 && e == hipSuccess; ++k) e = launch_dks(c, out.p, in.p, gauge, fmt, CL2_EVEN);
                if (e == hipSuccess) e = hipEventRecord(c->ev_start, c->stream);
                for (int k = 0; k < 5 && e == hipSuccess; ++k)
                    e = launch_dks(c, out.p, in.p, gauge, fmt, k & 1 ? CL2_ODD : CL2_EVEN);
                if (e == hipSuccess) e = hipEventRecord(c->ev_stop, c->stream);
                if (e == hipSuccess) e = hipEventSynchronize(c->ev_stop);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
                if (e != hipSuccess) break;
                if (ms / 5. < best) {
                    best = ms / 5.;
                    bt   = cl2_hop_tuning{zp, x, 1, 1, sh.first, sh.second};
                }
            }
    c->tune_manual = manual;
    if (e != hipSuccess) {
        c->tuned.erase(key);
        CL2_HIP(c, e);
    }
    c->tuned[key] = bt;
    if (best_ms) *best_ms = best;
    if (zplanes) *zplanes = bt.zplanes;
    if (xcd) *xcd = bt.xcd;
    if (block) *block = bt.block;
    if (occupancy) *occupancy = bt.occ;
    return CL2_OK;
}

/* ---------------- timing ---------------- */

int cl2_timer_start(cl2_ctx* c)
{
    if (!c) return CL2_E_ARG;
    CL2_HIP(c, hipEventRecord(c->ev_start, c->stream));
    return CL2_OK;
}

int cl2_timer_stop(cl2_ctx* c, double* ms)
{
    if (!c) return CL2_E_ARG;
    CL2_REQUIRE(c, ms != nullptr, "null output");
    CL2_HIP(c, hipEventRecord(c->ev_stop, c->stream));
    CL2_HIP(c, hipEventSynchronize(c->ev_stop));
    float f = 0.f;
    CL2_HIP(c, hipEventElapsedTime(&f, c->ev_start, c->ev_stop));
    *ms = f;
    return CL2_OK;
}

/* ---------------- byte / flop models ---------------- */

uint64_t cl2_reference_read_write_size(const cl2_ctx* c, const char* k)
{
    if (!c || !k) return 0;
    const uint64_t Seo = c->g.veo, D = 8, C = 2, R = 9;
    const std::string n(k);
    if (n == "dslash_eo" || n == "dslash_AND_M_tm_inverse_sitediagonal_eo" ||
        n == "dslash_AND_M_tm_inverse_sitediagonal_minus_eo")
        return (C * 12 * (2 * 4 + 1) + C * 2 * 4 * R) * D * Seo;  // fermions.cpp:860-871
    if (n == "gamma5_eo" || n.rfind("M_tm_", 0) == 0) return 48 * D * Seo;  // fermions.cpp:834-857,872-877
    if (n == "saxpy_AND_gamma5_eo") return C * D * Seo * (12 * (2 + 1) + 1);  // fermions.cpp:887-891
    if (n == "saxpy_eoprec") return C * D * Seo * (12 * (2 + 1) + 1);         // spinors.cpp:969-972
    if (n == "D_KS_eo") return (C * 3 * (2 * 4 + 1) + C * 2 * 4 * R) * D * Seo;  // fermions_staggered.cpp:189-193
    if (n == "saxpby_cplx_staggered_eoprec") return C * D * Seo * (3 * (2 + 1) + 2);  // spinors_staggered.cpp:1712-1715
    return 0;
}

uint64_t cl2_reference_flop_size(const cl2_ctx* c, const char* k)
{
    if (!c || !k) return 0;
    const uint64_t Seo = c->g.veo;
    const std::string n(k);
    const uint64_t dslash = 4 * 2 * (66 * 2 + 6 * 3 * 2 + 2 * 3 * 6);  // 1632, fermions.cpp:895-910
    const uint64_t diag   = 3 * 4 * 6;
    if (n == "dslash_eo") return Seo * dslash;
    if (n == "dslash_AND_M_tm_inverse_sitediagonal_eo" || n == "dslash_AND_M_tm_inverse_sitediagonal_minus_eo")
        return Seo * dslash + Seo * (diag + 3 * 4 * 2);
    if (n == "M_tm_sitediagonal" || n == "M_tm_sitediagonal_minus") return Seo * diag;
    if (n == "M_tm_inverse_sitediagonal" || n == "M_tm_inverse_sitediagonal_minus") return Seo * (diag + 24);
    if (n == "gamma5_eo") return Seo * 12;
    if (n == "M_tm_sitediagonal_AND_gamma5_eo" || n == "M_tm_sitediagonal_minus_AND_gamma5_eo")
        return Seo * diag + Seo * 12;                                // fermions.cpp:961-968
    if (n == "saxpy_AND_gamma5_eo") return Seo * 4 * 3 * (1 + 6 + 2);  // fermions.cpp:969-973
    if (n == "saxpy_eoprec") return Seo * 4 * 3 * (6 + 2);             // spinors.cpp:1085-1088
    if (n == "D_KS_eo") return Seo * 570;                              // fermions_staggered.cpp:259-263
    if (n == "saxpby_cplx_staggered_eoprec") return Seo * 3 * (2 * 6 + 2);  // spinors_staggered.cpp:1853-1856
    return 0;
}

uint64_t cl2_algorithmic_bytes(const cl2_ctx* c, const char* k, int fmt)
{
    if (!c || !k || !valid_fmt(fmt)) return 0;
    const uint64_t Seo = c->g.veo;
    const uint64_t L   = gauge_bytes(c, fmt) / (8 * Seo);
    const uint64_t S   = (fmt == CL2_GAUGE_F16_12) ? 52 : (fmt == CL2_GAUGE_F32_12) ? 96 : 192;
    const uint64_t s   = S / 4;
    const std::string n(k);
    if (n == "dslash_eo" || n == "dslash_AND_M_tm_inverse_sitediagonal_eo") return Seo * (8 * L + 2 * S);
    if (n == "dslash_AND_sitediagonal_saxpy_eo") return Seo * (8 * L + 3 * S);
    if (n == "D_KS_eo") return fmt == CL2_GAUGE_F16_12 ? 0 : Seo * (8 * L + 2 * s);
    if (n == "M_tm_sitediagonal" || n == "gamma5_eo" || n == "squarenorm") return Seo * 2 * S / (n == "squarenorm" ? 2 : 1);
    if (n == "saxpy" || n == "saxpy_AND_squarenorm") return Seo * 3 * S;
    if (n == "scalar_product") return Seo * 2 * S;
    if (n == "saxsbypz") return Seo * 4 * S;
    return 0;
}

}  // extern "C"
