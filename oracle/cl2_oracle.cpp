/*
 * CPU oracle for the CL2QCD fermion-inversion path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain FP64 restatement of the reference's kernels, operators and solvers,
 * written to follow the reference's arithmetic step by step so that results
 * agree to rounding.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  Each function cites
 * the reference file:line it restates.  Layouts are documented in cl2_oracle.h.
 */
#include "cl2_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int TDIR = 0, XDIR = 1, YDIR = 2, ZDIR = 3;  // operations_geometry.cl:72-75
const int EVEN = 0, ODD = 1;                       // common_header_files/globaldefs.hpp:38-39

struct C {
    double re, im;
};
struct V3 {
    C e[3];
};
struct Sp {
    V3 s[4];
};
struct M3 {
    C e[3][3];
};

int g_threads = 0;

inline C cmul(C a, C f)
{  // su3vec_times_complex component form, operations_su3vec.cl:91-101
    return C{a.re * f.re - a.im * f.im, a.im * f.re + a.re * f.im};
}
inline C cmul_conj(C a, C f)
{  // a * conj(f), operations_su3vec.cl:103-113
    return C{a.re * f.re + a.im * f.im, a.im * f.re - a.re * f.im};
}

inline V3 v_zero()
{
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = C{0., 0.};
    return r;
}
inline V3 v_acc(V3 a, V3 b)
{  // operations_su3vec.cl:168-178
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = C{a.e[i].re + b.e[i].re, a.e[i].im + b.e[i].im};
    return r;
}
inline V3 v_dim(V3 a, V3 b)
{  // a - b, operations_su3vec.cl:192-202
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = C{a.e[i].re - b.e[i].re, a.e[i].im - b.e[i].im};
    return r;
}
inline V3 v_acc_i(V3 a, V3 b)
{  // a + i*b, operations_su3vec.cl:180-190
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = C{a.e[i].re - b.e[i].im, a.e[i].im + b.e[i].re};
    return r;
}
inline V3 v_dim_i(V3 a, V3 b)
{  // a - i*b, operations_su3vec.cl:204-214
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = C{a.e[i].re + b.e[i].im, a.e[i].im - b.e[i].re};
    return r;
}
inline V3 v_times_c(V3 a, C f)
{
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = cmul(a.e[i], f);
    return r;
}
inline V3 v_times_c_conj(V3 a, C f)
{
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = cmul_conj(a.e[i], f);
    return r;
}
inline V3 v_times_r(V3 a, double f)
{
    V3 r;
    for (int i = 0; i < 3; ++i) r.e[i] = C{a.e[i].re * f, a.e[i].im * f};
    return r;
}

inline V3 su3_mul(const M3& u, V3 in)
{  // su3matrix_times_su3vec, operations_su3vec.cl:115-132
    V3 t;
    for (int r = 0; r < 3; ++r) {
        t.e[r].re = u.e[r][0].re * in.e[0].re + u.e[r][1].re * in.e[1].re + u.e[r][2].re * in.e[2].re -
                    u.e[r][0].im * in.e[0].im - u.e[r][1].im * in.e[1].im - u.e[r][2].im * in.e[2].im;
        t.e[r].im = u.e[r][0].re * in.e[0].im + u.e[r][1].re * in.e[1].im + u.e[r][2].re * in.e[2].im +
                    u.e[r][0].im * in.e[0].re + u.e[r][1].im * in.e[1].re + u.e[r][2].im * in.e[2].re;
    }
    return t;
}
inline V3 su3_dag_mul(const M3& u, V3 in)
{  // su3matrix_dagger_times_su3vec, operations_su3vec.cl:134-154
    V3 t;
    for (int r = 0; r < 3; ++r) {
        t.e[r].re = u.e[0][r].re * in.e[0].re + u.e[1][r].re * in.e[1].re + u.e[2][r].re * in.e[2].re +
                    u.e[0][r].im * in.e[0].im + u.e[1][r].im * in.e[1].im + u.e[2][r].im * in.e[2].im;
        t.e[r].im = u.e[0][r].re * in.e[0].im + u.e[1][r].re * in.e[1].im + u.e[2][r].re * in.e[2].im -
                    u.e[0][r].im * in.e[0].re - u.e[1][r].im * in.e[1].re - u.e[2][r].im * in.e[2].re;
    }
    return t;
}
inline M3 m_times_r(M3 u, double f)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) u.e[i][j] = C{u.e[i][j].re * f, u.e[i][j].im * f};
    return u;
}
inline M3 m_times_c(M3 u, C f)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) u.e[i][j] = cmul(u.e[i][j], f);
    return u;
}

inline Sp sp_zero()
{
    Sp r;
    for (int s = 0; s < 4; ++s) r.s[s] = v_zero();
    return r;
}
inline Sp sp_dim(const Sp& a, const Sp& b)
{
    Sp r;
    for (int s = 0; s < 4; ++s) r.s[s] = v_dim(a.s[s], b.s[s]);
    return r;
}
inline Sp sp_times_c(const Sp& a, C f)
{
    Sp r;
    for (int s = 0; s < 4; ++s) r.s[s] = v_times_c(a.s[s], f);
    return r;
}
inline Sp sp_times_r(const Sp& a, double f)
{
    Sp r;
    for (int s = 0; s < 4; ++s) r.s[s] = v_times_r(a.s[s], f);
    return r;
}
inline Sp gamma5_local(const Sp& in)
{  // fermionmatrix.cl:41-49
    Sp r = in;
    r.s[2] = v_times_r(in.s[2], -1.);
    r.s[3] = v_times_r(in.s[3], -1.);
    return r;
}
inline Sp m_diag_tm_local(const Sp& in, C f1, C f2)
{  // fermionmatrix.cl:28-36
    Sp r;
    r.s[0] = v_times_c(in.s[0], f1);
    r.s[1] = v_times_c(in.s[1], f1);
    r.s[2] = v_times_c(in.s[2], f2);
    r.s[3] = v_times_c(in.s[3], f2);
    return r;
}

inline Sp load_sp(const double* f, long i)
{
    Sp r;
    const double* p = f + 24 * i;
    for (int s = 0; s < 4; ++s)
        for (int c = 0; c < 3; ++c) r.s[s].e[c] = C{p[2 * (3 * s + c)], p[2 * (3 * s + c) + 1]};
    return r;
}
inline void store_sp(double* f, long i, const Sp& v)
{
    double* p = f + 24 * i;
    for (int s = 0; s < 4; ++s)
        for (int c = 0; c < 3; ++c) {
            p[2 * (3 * s + c)]     = v.s[s].e[c].re;
            p[2 * (3 * s + c) + 1] = v.s[s].e[c].im;
        }
}
inline V3 load_v3(const double* f, long i)
{
    V3 r;
    const double* p = f + 6 * i;
    for (int c = 0; c < 3; ++c) r.e[c] = C{p[2 * c], p[2 * c + 1]};
    return r;
}
inline void store_v3(double* f, long i, const V3& v)
{
    double* p = f + 6 * i;
    for (int c = 0; c < 3; ++c) {
        p[2 * c]     = v.e[c].re;
        p[2 * c + 1] = v.e[c].im;
    }
}
inline M3 load_link(const double* g, long link)
{
    M3 u;
    const double* p = g + 18 * link;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) u.e[r][c] = C{p[2 * (3 * r + c)], p[2 * (3 * r + c) + 1]};
    return u;
}

/* Lattice geometry, operations_geometry.cl:122-470 and the BC/chemical-potential
 * macros of hardware/code/opencl_module.cpp:110-122,190-203. */
struct Geo {
    int ns, nt;
    long vs, vol, veo;
    C bs, bt;  // exp(i*pi*theta/N)
    bool cpr, cpi;
    double expcpr, mexpcpr;
    C cpi_f;
    explicit Geo(const orc_lattice* l)
    {
        ns  = l->ns;
        nt  = l->nt;
        vs  = (long)ns * ns * ns;
        vol = vs * nt;
        veo = vol / 2;
        const double PI = 3.14159265358979323846;
        double as       = l->theta_s * PI / (double)ns;
        double at       = l->theta_t * PI / (double)nt;
        bs              = C{std::cos(as), std::sin(as)};
        bt              = C{std::cos(at), std::sin(at)};
        cpr             = l->use_cp_re != 0;
        cpi             = l->use_cp_im != 0;
        expcpr          = std::exp(l->cp_re);
        mexpcpr         = std::exp(-1. * l->cp_re);
        cpi_f           = C{std::cos(l->cp_im), std::sin(l->cp_im)};
    }
    inline void coords(long s, int& x, int& y, int& z, int& t) const
    {
        x      = (int)(s % ns);
        long r = s / ns;
        y      = (int)(r % ns);
        r /= ns;
        z = (int)(r % ns);
        t = (int)(r / ns);
    }
    inline long site(int x, int y, int z, int t) const
    {
        return x + (long)ns * (y + (long)ns * (z + (long)ns * t));
    }
    inline long up(long s, int dir) const
    {
        int x, y, z, t;
        coords(s, x, y, z, t);
        switch (dir) {
            case TDIR: t = (t + 1) % nt; break;
            case XDIR: x = (x + 1) % ns; break;
            case YDIR: y = (y + 1) % ns; break;
            default: z = (z + 1) % ns; break;
        }
        return site(x, y, z, t);
    }
    inline long dn(long s, int dir) const
    {
        int x, y, z, t;
        coords(s, x, y, z, t);
        switch (dir) {
            case TDIR: t = (t - 1 + nt) % nt; break;
            case XDIR: x = (x - 1 + ns) % ns; break;
            case YDIR: y = (y - 1 + ns) % ns; break;
            default: z = (z - 1 + ns) % ns; break;
        }
        return site(x, y, z, t);
    }
    inline int parity(long s) const
    {
        int x, y, z, t;
        coords(s, x, y, z, t);
        return (x + y + z + t) & 1;
    }
    /* get_even_st_idx / get_odd_st_idx (operations_geometry.cl:368-395): the
     * site of the given parity whose site index / 2 equals eo_idx. */
    inline long even_site(long i) const
    {
        long s = 2 * i;
        return parity(s) == 0 ? s : s + 1;
    }
    inline long odd_site(long i) const
    {
        long s = 2 * i;
        return parity(s) == 1 ? s : s + 1;
    }
};

template<typename F>
void site_loop(long n, F f)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) f(i);
}

/* dslash_eoprec_unified_local (ocl_kernel/fermionmatrix_eo.cl:30-253): the
 * kappa*BC-weighted hopping contributions of direction dir at site pos. */
Sp hop_eo(const Geo& g, const double* in, const double* gauge, long pos, int dir, double kappa)
{
    Sp out = sp_zero();
    V3 psi, phi;
    C bc = (dir == TDIR) ? C{kappa * g.bt.re, kappa * g.bt.im} : C{kappa * g.bs.re, kappa * g.bs.im};

    // mu = +dir
    long nb = g.up(pos, dir);
    Sp plus = load_sp(in, nb / 2);
    M3 U    = load_link(gauge, dir + 4 * pos);
    if (dir == XDIR) {
        psi      = v_acc_i(plus.s[0], plus.s[3]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[3] = v_dim_i(out.s[3], psi);
        psi      = v_acc_i(plus.s[1], plus.s[2]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[2] = v_dim_i(out.s[2], psi);
    } else if (dir == YDIR) {
        psi      = v_acc(plus.s[0], plus.s[3]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[3] = v_acc(out.s[3], psi);
        psi      = v_dim(plus.s[1], plus.s[2]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[2] = v_dim(out.s[2], psi);
    } else if (dir == ZDIR) {
        psi      = v_acc_i(plus.s[0], plus.s[2]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[2] = v_dim_i(out.s[2], psi);
        psi      = v_dim_i(plus.s[1], plus.s[3]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[3] = v_acc_i(out.s[3], psi);
    } else {
        if (g.cpr) U = m_times_r(U, g.expcpr);
        if (g.cpi) U = m_times_c(U, g.cpi_f);
        psi      = v_acc(plus.s[0], plus.s[2]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[2] = v_acc(out.s[2], psi);
        psi      = v_acc(plus.s[1], plus.s[3]);
        phi      = su3_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[3] = v_acc(out.s[3], psi);
    }

    // mu = -dir: U^dagger(x - mu), conjugate BC factor
    nb   = g.dn(pos, dir);
    plus = load_sp(in, nb / 2);
    U    = load_link(gauge, dir + 4 * nb);
    bc   = (dir == TDIR) ? C{kappa * g.bt.re, kappa * -g.bt.im} : C{kappa * g.bs.re, kappa * -g.bs.im};
    if (dir == XDIR) {
        psi      = v_dim_i(plus.s[0], plus.s[3]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[3] = v_acc_i(out.s[3], psi);
        psi      = v_dim_i(plus.s[1], plus.s[2]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[2] = v_acc_i(out.s[2], psi);
    } else if (dir == YDIR) {
        psi      = v_dim(plus.s[0], plus.s[3]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[3] = v_dim(out.s[3], psi);
        psi      = v_acc(plus.s[1], plus.s[2]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[2] = v_acc(out.s[2], psi);
    } else if (dir == ZDIR) {
        psi      = v_dim_i(plus.s[0], plus.s[2]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[2] = v_acc_i(out.s[2], psi);
        psi      = v_acc_i(plus.s[1], plus.s[3]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[3] = v_dim_i(out.s[3], psi);
    } else {
        if (g.cpr) U = m_times_r(U, g.mexpcpr);
        if (g.cpi) U = m_times_c(U, g.cpi_f);
        psi      = v_dim(plus.s[0], plus.s[2]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[0] = v_acc(out.s[0], psi);
        out.s[2] = v_dim(out.s[2], psi);
        psi      = v_dim(plus.s[1], plus.s[3]);
        phi      = su3_dag_mul(U, psi);
        psi      = v_times_c(phi, bc);
        out.s[1] = v_acc(out.s[1], psi);
        out.s[3] = v_dim(out.s[3], psi);
    }
    return out;
}

/* dslash_eo_for_site (ocl_kernel/fermionmatrix_eo_dslash.cl:28-48): -sum of hops, order T,X,Y,Z */
inline Sp dslash_site(const Geo& g, const double* in, const double* gauge, long pos, double kappa)
{
    Sp acc = sp_zero();
    acc    = sp_dim(acc, hop_eo(g, in, gauge, pos, TDIR, kappa));
    acc    = sp_dim(acc, hop_eo(g, in, gauge, pos, XDIR, kappa));
    acc    = sp_dim(acc, hop_eo(g, in, gauge, pos, YDIR, kappa));
    acc    = sp_dim(acc, hop_eo(g, in, gauge, pos, ZDIR, kappa));
    return acc;
}

/* get_modified_stagg_phase (ocl_kernel/operations_staggered.cl:93-104): eta_mu(x) * BC phase */
inline C stagg_phase(const Geo& g, long pos, int dir)
{
    int x, y, z, t;
    g.coords(pos, x, y, z, t);
    double eta;
    C b = g.bs;
    switch (dir) {
        case XDIR: eta = 1.; break;
        case YDIR: eta = 1 - 2. * (x % 2); break;
        case ZDIR: eta = 1 - 2. * ((x + y) % 2); break;
        default:
            eta = 1 - 2. * ((x + y + z) % 2);
            b   = g.bt;
            break;
    }
    return C{eta * b.re, eta * b.im};
}

/* D_KS_eo_local (ocl_kernel/fermionmatrix_staggered_eo_DKS_local.cl:65-136) */
inline V3 dks_dir(const Geo& g, const double* in, const double* gauge, long pos, int dir)
{
    V3 out = v_zero();
    long nb = g.up(pos, dir);
    V3 plus = load_v3(in, nb / 2);
    M3 U    = load_link(gauge, dir + 4 * pos);
    if (g.cpi && dir == TDIR) U = m_times_c(U, g.cpi_f);
    V3 chi = su3_mul(U, plus);
    C eta  = stagg_phase(g, pos, dir);
    eta.re *= 0.5;
    eta.im *= 0.5;
    chi = v_times_c(chi, eta);
    out = v_acc(out, chi);

    nb   = g.dn(pos, dir);
    plus = load_v3(in, nb / 2);
    U    = load_link(gauge, dir + 4 * nb);
    if (g.cpi && dir == TDIR) U = m_times_c(U, g.cpi_f);
    chi = su3_dag_mul(U, plus);
    eta = stagg_phase(g, pos, dir);
    eta.re *= 0.5;
    eta.im *= 0.5;
    chi = v_times_c_conj(chi, eta);
    out = v_dim(out, chi);
    return out;
}

/* ---- eo spinor BLAS helpers on flat double arrays ---- */
void blas_copy(double* out, const double* in, long n, int ncomp)
{
    std::memcpy(out, in, sizeof(double) * 2 * ncomp * n);
}

struct Cplx {
    double re, im;
};
Cplx cdiv(Cplx n, Cplx d)
{  // complexdivide, common_header_files/operations_complex.hpp:59-66
    double norm = d.re * d.re + d.im * d.im;
    return Cplx{(n.re * d.re + n.im * d.im) / norm, (n.im * d.re - n.re * d.im) / norm};
}
Cplx cmulc(Cplx a, Cplx b)
{
    return Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

/* Wilson eo operator state: the temporaries the reference operator classes own
 * (fermionmatrix.hpp:207-208,309). */
struct EoOp {
    const orc_lattice* lat;
    const double* gauge;
    int action;
    double kappa, mubar;
    int merged;
    long veo;
    std::vector<double> tmp, tmp2, tmpq;
    EoOp(const orc_lattice* l, const double* g, int a, double k, double m, int mg)
        : lat(l), gauge(g), action(a), kappa(k), mubar(m), merged(mg)
    {
        veo = orc_veo(l);
        tmp.assign(24 * veo, 0.);
        tmp2.assign(24 * veo, 0.);
        tmpq.assign(24 * veo, 0.);
    }
    /* Aee family, fermionmatrix.cpp:224-591 */
    void aee(double* out, const double* in, bool minus, bool g5)
    {
        if (action == ORC_ACTION_WILSON) {
            orc_dslash_eo(lat, tmp.data(), in, gauge, ODD, kappa);
            orc_dslash_eo(lat, out, tmp.data(), gauge, EVEN, kappa);
            if (g5)
                orc_saxpy_and_gamma5_eo(lat, out, out, in, 1., 0.);
            else
                orc_saxpy(out, out, in, 1., 0., veo, 12);
        } else {
            orc_dslash_eo(lat, tmp.data(), in, gauge, ODD, kappa);
            orc_m_tm(lat, tmp2.data(), tmp.data(), mubar, minus ? ORC_MTM_INVERSE_MINUS : ORC_MTM_INVERSE);
            orc_dslash_eo(lat, out, tmp2.data(), gauge, EVEN, kappa);
            orc_m_tm(lat, tmp.data(), in, mubar, minus ? ORC_MTM_SITEDIAG_MINUS : ORC_MTM_SITEDIAG);
            if (g5)
                orc_saxpy_and_gamma5_eo(lat, out, out, tmp.data(), 1., 0.);
            else
                orc_saxpy(out, out, tmp.data(), 1., 0., veo, 12);
        }
    }
    /* Qplus_eo / Qminus_eo, fermionmatrix.cpp:592-624 */
    void q(double* out, const double* in, bool minus)
    {
        if (!merged) {
            aee(out, in, minus, false);
            orc_gamma5_eo(lat, out);
        } else {
            aee(out, in, minus, true);
        }
    }
    void apply(int op, double* out, const double* in)
    {
        switch (op) {
            case ORC_OP_AEE: aee(out, in, false, false); break;
            case ORC_OP_AEE_AND_GAMMA5: aee(out, in, false, true); break;
            case ORC_OP_AEE_MINUS: aee(out, in, true, false); break;
            case ORC_OP_AEE_MINUS_AND_GAMMA5: aee(out, in, true, true); break;
            case ORC_OP_QPLUS_EO: q(out, in, false); break;
            case ORC_OP_QMINUS_EO: q(out, in, true); break;
            default:  // QplusQminus_eo, fermionmatrix.cpp:658-664
                q(tmpq.data(), in, true);
                q(out, tmpq.data(), false);
                break;
        }
    }
};

}  // namespace

extern "C" {

void orc_set_threads(int n)
{
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}
int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

long orc_volume(const orc_lattice* lat)
{
    return (long)lat->ns * lat->ns * lat->ns * lat->nt;
}
long orc_veo(const orc_lattice* lat)
{
    return orc_volume(lat) / 2;
}
int orc_even_site(const orc_lattice* lat, long eo_idx, long* site)
{
    Geo g(lat);
    *site = g.even_site(eo_idx);
    return 0;
}
int orc_odd_site(const orc_lattice* lat, long eo_idx, long* site)
{
    Geo g(lat);
    *site = g.odd_site(eo_idx);
    return 0;
}

/* fill<>() of meta/type_ops.cpp:29-79: srand48(seed) then drand48() for every
 * real in storage order, which for every basetype is a flat (re, im) sweep. */
void orc_fill_doubles_drand48(double* out, long n, int seed)
{
    srand48(seed);
    for (long i = 0; i < n; ++i) out[i] = drand48();
}

/* cold start: unit matrix on every link */
void orc_gauge_cold(const orc_lattice* lat, double* gauge)
{
    long nl = 4 * orc_volume(lat);
    std::memset(gauge, 0, sizeof(double) * 18 * nl);
    for (long l = 0; l < nl; ++l) {
        gauge[18 * l + 0]  = 1.;
        gauge[18 * l + 8]  = 1.;
        gauge[18 * l + 16] = 1.;
    }
}

/* ILDG/LIME reader, ildg_io/ildgIo_gaugefield.cpp:142-221: file order is
 * [t][z][y][x][mu_ildg][row][col][re,im] big endian; site order equals the
 * internal site index and mu_internal = (mu_ildg + 1) % 4. */
int orc_ildg_read(const char* path, const orc_lattice* lat, double* gauge, char* err, int errlen)
{
    auto fail = [&](const std::string& m) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", m.c_str());
        return ORC_ERR_ARG;
    };
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(std::string("cannot open ") + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[65536];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    std::fclose(f);
    auto be64 = [&](size_t off) {
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v = (v << 8) | buf[off + i];
        return v;
    };
    auto be32 = [&](size_t off) {
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) v = (v << 8) | buf[off + i];
        return v;
    };
    size_t off = 0;
    size_t data_off = 0, data_len = 0;
    int precision = 64;
    while (off + 144 <= buf.size()) {
        if (be32(off) != 0x456789abu) return fail("bad LIME magic");
        uint64_t len = be64(off + 8);
        std::string type(reinterpret_cast<const char*>(&buf[off + 16]), strnlen((const char*)&buf[off + 16], 128));
        size_t payload = off + 144;
        if (payload + len > buf.size()) return fail("truncated LIME record");
        if (type == "ildg-format") {
            std::string xml(reinterpret_cast<const char*>(&buf[payload]), len);
            size_t p = xml.find("<precision>");
            if (p != std::string::npos) precision = std::atoi(xml.c_str() + p + 11);
            auto dim = [&](const char* tag) {
                size_t q = xml.find(tag);
                return q == std::string::npos ? -1 : std::atoi(xml.c_str() + q + std::strlen(tag));
            };
            int lx = dim("<lx>"), ly = dim("<ly>"), lz = dim("<lz>"), lt = dim("<lt>");
            if (lx != lat->ns || ly != lat->ns || lz != lat->ns || lt != lat->nt)
                return fail("lattice extents in file do not match");
        } else if (type == "ildg-binary-data") {
            data_off = payload;
            data_len = len;
        }
        off = payload + ((len + 7) / 8) * 8;
    }
    long vol = orc_volume(lat);
    size_t nreal = (size_t)vol * 4 * 18;
    size_t esz   = precision == 32 ? 4 : 8;
    if (data_len != nreal * esz) return fail("ildg-binary-data has wrong size");
    for (long s = 0; s < vol; ++s)
        for (int l = 0; l < 4; ++l)
            for (int k = 0; k < 18; ++k) {
                size_t src = data_off + ((size_t)(s * 4 + l) * 18 + k) * esz;
                double v;
                if (esz == 8) {
                    uint64_t b = be64(src);
                    std::memcpy(&v, &b, 8);
                } else {
                    uint32_t b = be32(src);
                    float fv;
                    std::memcpy(&fv, &b, 4);
                    v = fv;
                }
                gauge[(size_t)((l + 1) % 4 + 4 * s) * 18 + k] = v;
            }
    return ORC_OK;
}

/* average plaquette normalised to 1 for the unit field: (1/(6*3*V)) sum Re tr P */
double orc_plaquette(const orc_lattice* lat, const double* gauge)
{
    Geo g(lat);
    double sum = 0.;
    for (long s = 0; s < g.vol; ++s)
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                M3 a = load_link(gauge, mu + 4 * s), b = load_link(gauge, nu + 4 * g.up(s, mu));
                M3 c = load_link(gauge, mu + 4 * g.up(s, nu)), d = load_link(gauge, nu + 4 * s);
                // Re tr (a b c^+ d^+)
                C ab[3][3], abc[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        C r{0, 0};
                        for (int k = 0; k < 3; ++k) {
                            C p = cmul(a.e[i][k], b.e[k][j]);
                            r.re += p.re;
                            r.im += p.im;
                        }
                        ab[i][j] = r;
                    }
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        C r{0, 0};
                        for (int k = 0; k < 3; ++k) {
                            C p = cmul_conj(ab[i][k], c.e[j][k]);
                            r.re += p.re;
                            r.im += p.im;
                        }
                        abc[i][j] = r;
                    }
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) sum += cmul_conj(abc[i][k], d.e[i][k]).re;
            }
    return sum / (6. * 3. * (double)g.vol);
}

void orc_convert_to_eo(const orc_lattice* lat, const double* full, double* even, double* odd)
{
    Geo g(lat);
    for (long n = 0; n < g.veo; ++n) {
        std::memcpy(even + 24 * n, full + 24 * g.even_site(n), 24 * sizeof(double));
        std::memcpy(odd + 24 * n, full + 24 * g.odd_site(n), 24 * sizeof(double));
    }
}
void orc_convert_from_eo(const orc_lattice* lat, const double* even, const double* odd, double* full)
{
    Geo g(lat);
    for (long n = 0; n < g.veo; ++n) {
        std::memcpy(full + 24 * g.even_site(n), even + 24 * n, 24 * sizeof(double));
        std::memcpy(full + 24 * g.odd_site(n), odd + 24 * n, 24 * sizeof(double));
    }
}

/* __kernel dslash_eo, fermionmatrix_eo_dslash.cl:50-58: evenodd == ODD writes
 * the even sites (reading odd neighbours), EVEN writes the odd sites. */
void orc_dslash_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd,
                   double kappa)
{
    Geo g(lat);
    site_loop(g.veo, [&](long i) {
        long pos = (evenodd == ODD) ? g.even_site(i) : g.odd_site(i);
        store_sp(out, pos / 2, dslash_site(g, in, gauge, pos, kappa));
    });
}

/* orc_dslash_eo restricted to a list of output eo indices (the same site formula,
 * fermionmatrix_eo_dslash.cl:50-58): out[k] is the output spinor at eo index eo[k].
 * Sampled checks of lattices too large for a full CPU pass use it. */
void orc_dslash_eo_sites(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd,
                         double kappa, const long* eo, long n)
{
    Geo g(lat);
    site_loop(n, [&](long k) {
        long pos = (evenodd == ODD) ? g.even_site(eo[k]) : g.odd_site(eo[k]);
        store_sp(out, k, dslash_site(g, in, gauge, pos, kappa));
    });
}

/* fermionmatrix_eo_m.cl:22-109 and fermionmatrix_eo_m_merged.cl:22-64 */
void orc_m_tm(const orc_lattice* lat, double* out, const double* in, double mubar, int variant)
{
    long veo = orc_veo(lat);
    C f1, f2;
    bool inv = false;
    switch (variant) {
        case ORC_MTM_SITEDIAG: f1 = C{1., mubar}, f2 = C{1., -1. * mubar}; break;
        case ORC_MTM_INVERSE: f1 = C{1., -1. * mubar}, f2 = C{1., mubar}, inv = true; break;
        case ORC_MTM_SITEDIAG_MINUS: f1 = C{1., -1. * mubar}, f2 = C{1., mubar}; break;
        case ORC_MTM_INVERSE_MINUS: f1 = C{1., mubar}, f2 = C{1., -1. * mubar}, inv = true; break;
        case ORC_MTM_SITEDIAG_AND_GAMMA5: f1 = C{1., mubar}, f2 = C{-1., 1. * mubar}; break;
        default: f1 = C{1., -mubar}, f2 = C{-1., -1. * mubar}; break;  // SITEDIAG_MINUS_AND_GAMMA5
    }
    double denom = 1. / (1. + mubar * mubar);
    site_loop(veo, [&](long i) {
        Sp r = m_diag_tm_local(load_sp(in, i), f1, f2);
        if (inv) r = sp_times_r(r, denom);
        store_sp(out, i, r);
    });
}

/* dslash_AND_M_tm_inverse_sitediagonal{,_minus}_eo,
 * fermionmatrix_eo_dslash_AND_M_tm_inverse_sitediagonal{,_minus}.cl:30-62 */
void orc_dslash_minv_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd,
                        double kappa, double mubar, int minus)
{
    Geo g(lat);
    C f1 = minus ? C{1., mubar} : C{1., -1. * mubar};
    C f2 = minus ? C{1., -1. * mubar} : C{1., mubar};
    site_loop(g.veo, [&](long i) {
        long pos     = (evenodd == ODD) ? g.even_site(i) : g.odd_site(i);
        Sp acc       = dslash_site(g, in, gauge, pos, kappa);
        Sp t         = m_diag_tm_local(acc, f1, f2);
        double denom = 1. / (1. + mubar * mubar);
        store_sp(out, pos / 2, sp_times_r(t, denom));
    });
}

void orc_gamma5_eo(const orc_lattice* lat, double* inout)
{  // fermionmatrix_eo_gamma5.cl:22-33
    site_loop(orc_veo(lat), [&](long i) { store_sp(inout, i, gamma5_local(load_sp(inout, i))); });
}

void orc_saxpy_and_gamma5_eo(const orc_lattice* lat, double* out, const double* x, const double* y, double are,
                             double aim)
{  // fermionmatrix_saxpy_AND_gamma5_eo.cl:23-38: gamma5 (y - alpha x)
    C a{are, aim};
    site_loop(orc_veo(lat), [&](long i) {
        Sp xt = sp_times_c(load_sp(x, i), a);
        xt    = sp_dim(load_sp(y, i), xt);
        store_sp(out, i, gamma5_local(xt));
    });
}

/* Non-eo M_wilson / M_tm_plus / M_tm_minus (fermionmatrix_m.cl:28-62,
 * fermionmatrix_m_tm_{plus,minus}.cl): diag(in) - sum of hops over all sites.
 * Restated on top of the eo hop so that the reference's non-eo goldens
 * cross-check the hopping term. */
void orc_m_full(const orc_lattice* lat, double* out, const double* in, const double* gauge, double kappa,
                double mubar, int action, int minus)
{
    Geo g(lat);
    std::vector<double> ev(24 * g.veo), od(24 * g.veo), oe(24 * g.veo), oo(24 * g.veo);
    orc_convert_to_eo(lat, in, ev.data(), od.data());
    C f1 = minus ? C{1., -1. * mubar} : C{1., mubar};
    C f2 = minus ? C{1., mubar} : C{1., -1. * mubar};
    for (int par = 0; par < 2; ++par) {
        const double* src = par == 0 ? od.data() : ev.data();  // neighbours of the opposite parity
        const double* dia = par == 0 ? ev.data() : od.data();
        double* dst       = par == 0 ? oe.data() : oo.data();
        site_loop(g.veo, [&](long i) {
            long pos = par == 0 ? g.even_site(i) : g.odd_site(i);
            Sp d     = load_sp(dia, i);
            Sp acc   = (action == ORC_ACTION_TWISTEDMASS) ? m_diag_tm_local(d, f1, f2) : d;
            acc      = sp_dim(acc, hop_eo(g, src, gauge, pos, TDIR, kappa));
            acc      = sp_dim(acc, hop_eo(g, src, gauge, pos, XDIR, kappa));
            acc      = sp_dim(acc, hop_eo(g, src, gauge, pos, YDIR, kappa));
            acc      = sp_dim(acc, hop_eo(g, src, gauge, pos, ZDIR, kappa));
            store_sp(dst, i, acc);
        });
    }
    orc_convert_from_eo(lat, oe.data(), oo.data(), out);
}

int orc_apply_eo(const orc_lattice* lat, int op, double* out, const double* in, const double* gauge, int action,
                 double kappa, double mubar, int merged)
{
    if (op < ORC_OP_AEE || op > ORC_OP_QPLUSQMINUS_EO) return ORC_ERR_ARG;
    EoOp A(lat, gauge, action, kappa, mubar, merged);
    A.apply(op, out, in);
    return ORC_OK;
}

/* ---- BLAS (spinorfield_eo_*.cl, spinorfield_staggered_eo_*.cl) ----
 * ncomp = complex components per site (12 Wilson, 3 staggered). */
void orc_saxpy(double* out, const double* x, const double* y, double are, double aim, long n, int ncomp)
{  // Wilson saxpy: out = y - alpha*x (spinorfield_eo_saxpy.cl:21-35)
    long m = n * ncomp;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < m; ++k) {
        C xt{x[2 * k], x[2 * k + 1]};
        xt = cmul(xt, C{are, aim});
        out[2 * k]     = y[2 * k] - xt.re;
        out[2 * k + 1] = y[2 * k + 1] - xt.im;
    }
}
void orc_sax(double* out, const double* x, double are, double aim, long n, int ncomp)
{  // spinorfield_eo_sax.cl:23-35
    long m = n * ncomp;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < m; ++k) {
        C r            = cmul(C{x[2 * k], x[2 * k + 1]}, C{are, aim});
        out[2 * k]     = r.re;
        out[2 * k + 1] = r.im;
    }
}
void orc_saxsbypz(double* out, const double* x, const double* y, const double* z, double are, double aim, double bre,
                  double bim, long n, int ncomp)
{  // alpha*x + beta*y + z, summed as (beta y) + (alpha x) + z (spinorfield_eo_saxsbypz.cl:22-43)
    long m = n * ncomp;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < m; ++k) {
        C xt = cmul(C{x[2 * k], x[2 * k + 1]}, C{are, aim});
        C yt = cmul(C{y[2 * k], y[2 * k + 1]}, C{bre, bim});
        out[2 * k]     = yt.re + xt.re + z[2 * k];
        out[2 * k + 1] = yt.im + xt.im + z[2 * k + 1];
    }
}
void orc_dot(const double* x, const double* y, long n, int ncomp, double* re, double* im)
{  // sum conj(x) y (operations_su3vec.cl:40-48)
    long m = n * ncomp;
    double sr = 0., si = 0.;
#pragma omp parallel for reduction(+ : sr, si) schedule(static)
    for (long k = 0; k < m; ++k) {
        sr += x[2 * k] * y[2 * k] + x[2 * k + 1] * y[2 * k + 1];
        si += x[2 * k] * y[2 * k + 1] - y[2 * k] * x[2 * k + 1];
    }
    *re = sr;
    *im = si;
}
double orc_norm2(const double* x, long n, int ncomp)
{
    long m = 2 * n * ncomp;
    double s = 0.;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long k = 0; k < m; ++k) s += x[k] * x[k];
    return s;
}

/* __kernel D_KS_eo (fermionmatrix_staggered_eo_DKS.cl:50-69): output parity == evenodd */
void orc_dks_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, int evenodd)
{
    Geo g(lat);
    site_loop(g.veo, [&](long i) {
        long pos = (evenodd == EVEN) ? g.even_site(i) : g.odd_site(i);
        V3 acc   = v_zero();
        for (int dir = 0; dir < 4; ++dir) acc = v_acc(acc, dks_dir(g, in, gauge, pos, dir));
        store_v3(out, pos / 2, acc);
    });
}

/* staggered saxpby: alpha*x + beta*y (spinorfield_staggered_eo_saxpby_cplx.cl) */
void orc_stagg_saxpby(double* out, const double* x, const double* y, double are, double aim, double bre, double bim,
                      long n)
{
    long m = 3 * n;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < m; ++k) {
        C xt = cmul(C{x[2 * k], x[2 * k + 1]}, C{are, aim});
        C yt = cmul(C{y[2 * k], y[2 * k + 1]}, C{bre, bim});
        out[2 * k]     = xt.re + yt.re;
        out[2 * k + 1] = xt.im + yt.im;
    }
}

/* MdagM_eo (physics/fermionmatrix/fermionmatrix_stagg.cpp:72-91) */
void orc_mdagm_eo(const orc_lattice* lat, double* out, const double* in, const double* gauge, double mass,
                  int upper_left)
{
    long veo = orc_veo(lat);
    std::vector<double> tmp(6 * veo);
    if (upper_left == EVEN) {
        orc_dks_eo(lat, tmp.data(), in, gauge, ODD);
        orc_dks_eo(lat, out, tmp.data(), gauge, EVEN);
    } else {
        orc_dks_eo(lat, tmp.data(), in, gauge, EVEN);
        orc_dks_eo(lat, out, tmp.data(), gauge, ODD);
    }
    orc_stagg_saxpby(out, in, out, mass * mass, 0., -1., 0., veo);
}

/* cg_singledev, physics/algorithms/solvers/cg.cpp:277-400 */
int orc_cg_eo(const orc_lattice* lat, int op, const double* gauge, int action, double kappa, double mubar,
              int merged_fermion, double* x, const double* b, double prec, int cgmax, int iter_refresh,
              int check_every, int merged_spinor, int* iters)
{
    (void)merged_spinor;  // saxpy_AND_squarenorm and saxpy + scalar_product give the same numbers
    long veo = orc_veo(lat);
    EoOp A(lat, gauge, action, kappa, mubar, merged_fermion);
    std::vector<double> p(24 * veo), rn(24 * veo), v(24 * veo);
    Cplx omega{0, 0}, rho{0, 0}, rho_next{0, 0}, alpha, beta;
    int iter;
    for (iter = 0; iter < cgmax; ++iter) {
        if (iter % iter_refresh == 0) {
            A.apply(op, rn.data(), x);
            orc_saxpy(rn.data(), rn.data(), b, 1., 0., veo, 12);  // rn = b - A x
            blas_copy(p.data(), rn.data(), veo, 12);
            orc_dot(rn.data(), rn.data(), veo, 12, &omega.re, &omega.im);
        } else {
            omega = rho_next;
        }
        A.apply(op, v.data(), p.data());
        orc_dot(p.data(), v.data(), veo, 12, &rho.re, &rho.im);
        alpha     = cdiv(omega, rho);
        Cplx tmp1 = cmulc(Cplx{-1., 0.}, alpha);
        orc_saxpy(x, p.data(), x, tmp1.re, tmp1.im, veo, 12);              // x = x + alpha p
        orc_saxpy(rn.data(), v.data(), rn.data(), alpha.re, alpha.im, veo, 12);  // rn = rn - alpha v
        orc_dot(rn.data(), rn.data(), veo, 12, &rho_next.re, &rho_next.im);
        if (iter % check_every == 0) {
            double resid = rho_next.re;
            if (resid != resid) {
                *iters = iter;
                return ORC_SOLVER_STUCK;
            }
            if (resid < prec) {
                *iters = iter;
                return ORC_OK;
            }
        }
        beta      = cdiv(rho_next, omega);
        Cplx tmp2 = cmulc(Cplx{-1., 0.}, beta);
        orc_saxpy(p.data(), p.data(), rn.data(), tmp2.re, tmp2.im, veo, 12);  // p = rn + beta p
    }
    *iters = iter;
    return ORC_SOLVER_DID_NOT_SOLVE;
}

/* bicgstab_fast / bicgstab_save, physics/algorithms/solvers/bicgstab.cpp:401-733 */
int orc_bicgstab_eo(const orc_lattice* lat, int op, const double* gauge, int action, double kappa, double mubar,
                    int merged_fermion, double* x, const double* b, double prec, int cgmax, int iter_refresh,
                    int save_variant, int* iters)
{
    long veo = orc_veo(lat);
    EoOp A(lat, gauge, action, kappa, mubar, merged_fermion);
    std::vector<double> p(24 * veo), rn(24 * veo), rhat(24 * veo), s(24 * veo), t(24 * veo), v(24 * veo),
        aux(24 * veo);
    Cplx alpha{1, 0}, beta{0, 0}, omega{1, 0}, rho{1, 0}, rho_next{0, 0}, tmp1, tmp2;
    const Cplx one{1., 0.}, minus_one{-1., 0.};
    int iter;
    double resid;
    if (save_variant) {
        for (iter = 0; iter < cgmax; ++iter) {
            if (iter % iter_refresh == 0) {
                std::fill(v.begin(), v.end(), 0.);
                std::fill(p.begin(), p.end(), 0.);
                A.apply(op, rn.data(), x);
                orc_saxpy(rn.data(), rn.data(), b, one.re, one.im, veo, 12);
                blas_copy(rhat.data(), rn.data(), veo, 12);
                alpha = one;
                omega = one;
                rho   = one;
            }
            orc_dot(rhat.data(), rn.data(), veo, 12, &rho_next.re, &rho_next.im);
            if (std::fabs(rho_next.re) < 1e-25 && std::fabs(rho_next.im) < 1e-25) {
                *iters = iter;
                return ORC_SOLVER_STUCK;
            }
            tmp1 = cdiv(rho_next, rho);
            rho  = rho_next;
            tmp2 = cdiv(alpha, omega);
            beta = cmulc(tmp1, tmp2);
            tmp1 = cmulc(beta, omega);
            tmp2 = cmulc(minus_one, tmp1);
            orc_saxsbypz(p.data(), p.data(), v.data(), rn.data(), beta.re, beta.im, tmp2.re, tmp2.im, veo, 12);
            A.apply(op, v.data(), p.data());
            orc_dot(rhat.data(), v.data(), veo, 12, &tmp1.re, &tmp1.im);
            alpha = cdiv(rho, tmp1);
            orc_saxpy(s.data(), v.data(), rn.data(), alpha.re, alpha.im, veo, 12);
            A.apply(op, t.data(), s.data());
            orc_dot(t.data(), s.data(), veo, 12, &tmp1.re, &tmp1.im);
            orc_dot(t.data(), t.data(), veo, 12, &tmp2.re, &tmp2.im);
            omega = cdiv(tmp1, tmp2);
            orc_saxpy(rn.data(), t.data(), s.data(), omega.re, omega.im, veo, 12);
            orc_saxsbypz(x, p.data(), s.data(), x, alpha.re, alpha.im, omega.re, omega.im, veo, 12);
            resid = orc_norm2(rn.data(), veo, 12);
            if (resid != resid) {
                *iters = iter;
                return ORC_SOLVER_STUCK;
            }
            if (resid < prec) {
                A.apply(op, aux.data(), x);
                orc_saxpy(aux.data(), aux.data(), b, one.re, one.im, veo, 12);
                double trueresid = orc_norm2(aux.data(), veo, 12);
                if (trueresid < prec) {
                    *iters = iter;
                    return ORC_OK;
                }
            }
        }
        *iters = iter;
        return ORC_SOLVER_DID_NOT_SOLVE;
    }
    for (iter = 0; iter < cgmax; ++iter) {
        if (iter % iter_refresh == 0) {
            A.apply(op, rn.data(), x);
            orc_saxpy(p.data(), rn.data(), b, one.re, one.im, veo, 12);
            blas_copy(rhat.data(), p.data(), veo, 12);
            blas_copy(rn.data(), p.data(), veo, 12);
            orc_dot(rhat.data(), rn.data(), veo, 12, &rho.re, &rho.im);
        }
        resid = orc_norm2(rn.data(), veo, 12);
        if (resid != resid) {
            *iters = iter;
            return ORC_SOLVER_STUCK;
        }
        if (resid < prec) {
            *iters = iter;
            return ORC_OK;
        }
        A.apply(op, v.data(), p.data());
        orc_dot(rhat.data(), v.data(), veo, 12, &tmp1.re, &tmp1.im);
        alpha = cdiv(rho, tmp1);
        orc_saxpy(s.data(), v.data(), rn.data(), alpha.re, alpha.im, veo, 12);
        A.apply(op, t.data(), s.data());
        orc_dot(t.data(), s.data(), veo, 12, &tmp1.re, &tmp1.im);
        orc_dot(t.data(), t.data(), veo, 12, &tmp2.re, &tmp2.im);
        omega = cdiv(tmp1, tmp2);
        orc_saxsbypz(x, p.data(), s.data(), x, alpha.re, alpha.im, omega.re, omega.im, veo, 12);
        orc_saxpy(rn.data(), t.data(), s.data(), omega.re, omega.im, veo, 12);
        orc_dot(rhat.data(), rn.data(), veo, 12, &rho_next.re, &rho_next.im);
        if (std::fabs(rho_next.re) < 1e-25 && std::fabs(rho_next.im) < 1e-25) {
            *iters = iter;
            return ORC_SOLVER_STUCK;
        }
        tmp1 = cdiv(rho_next, rho);
        tmp2 = cdiv(alpha, omega);
        beta = cmulc(tmp1, tmp2);
        tmp1 = cmulc(beta, omega);
        tmp2 = cmulc(minus_one, tmp1);
        orc_saxsbypz(p.data(), p.data(), v.data(), rn.data(), beta.re, beta.im, tmp2.re, tmp2.im, veo, 12);
        rho = rho_next;
    }
    *iters = iter;
    return ORC_SOLVER_DID_NOT_SOLVE;
}

/* cg_m / SolverShifted::solve, physics/algorithms/solver_shifted.cpp:114-475,
 * recurrences ocl_kernel/operations_real.cl:20-39 with the argument mapping of
 * physics/lattices/algebra_real.cpp:144-200. */
int orc_cgm(const orc_lattice* lat, const double* gauge, double mass, int upper_left, double* const* x,
            const double* b, const double* sigma, int nshift, double prec, int cgmax, int check_every,
            int merged_spinor, int* iters)
{
    (void)merged_spinor;  // sax_vec_and_squarenorm and sax + squarenorm give the same numbers
    long veo = orc_veo(lat);
    long nd  = 6 * veo;
    if (orc_norm2(b, veo, 3) == 0.) {
        for (int k = 0; k < nshift; ++k) std::memset(x[k], 0, sizeof(double) * nd);
        *iters = 0;
        return ORC_OK;
    }
    std::vector<double> r(b, b + nd), p(b, b + nd), v(nd);
    std::vector<std::vector<double>> ps(nshift, std::vector<double>(b, b + nd));
    std::vector<double> zeta_prev(nshift, 1.), zeta(nshift, 1.), zeta_foll(nshift), alpha_vec(nshift, 0.),
        beta_vec(nshift);
    std::vector<char> conv(nshift, 0);
    int nconv = 0;
    for (int k = 0; k < nshift; ++k) std::memset(x[k], 0, sizeof(double) * nd);
    double im_unused;
    double tmp1, tmp2, tmp3;
    orc_dot(r.data(), r.data(), veo, 3, &tmp1, &im_unused);
    double beta_scalar = 1.0, alpha_scalar = 0.0, beta_scalar_prev, alpha_scalar_prev;
    int it = 0;
    while (it < cgmax) {
        beta_scalar_prev = beta_scalar;
        orc_mdagm_eo(lat, v.data(), p.data(), gauge, mass, upper_left);
        orc_dot(p.data(), v.data(), veo, 3, &tmp3, &im_unused);
        beta_scalar = tmp1 / tmp3;
        beta_scalar = 0.0 - beta_scalar;
        orc_stagg_saxpby(r.data(), v.data(), r.data(), beta_scalar, 0., 1., 0., veo);  // r = beta v + r
        orc_dot(r.data(), r.data(), veo, 3, &tmp2, &im_unused);
        alpha_scalar_prev = alpha_scalar;
        alpha_scalar      = tmp2 / tmp1;
        orc_stagg_saxpby(p.data(), p.data(), r.data(), alpha_scalar, 0., 1., 0., veo);  // p = alpha p + r
        for (int k = 0; k < nshift; ++k) {
            double a = zeta[k], bb = zeta_prev[k], c = beta_scalar_prev, d = beta_scalar, e = alpha_scalar_prev,
                   f = sigma[k];
            zeta_foll[k] = (a * bb * c) / (d * e * (bb - a) + bb * c * (1 - f * d));
            beta_vec[k]  = -beta_scalar * zeta_foll[k] / zeta[k];
            alpha_vec[k] = -alpha_scalar * zeta_foll[k] * beta_vec[k] / (zeta[k] * beta_scalar);
        }
        for (int k = 0; k < nshift; ++k) {
            if (conv[k]) continue;
            orc_stagg_saxpby(x[k], ps[k].data(), x[k], beta_vec[k], 0., 1., 0., veo);
            orc_stagg_saxpby(ps[k].data(), r.data(), ps[k].data(), zeta_foll[k], 0., alpha_vec[k], 0., veo);
            if (it % check_every == 0) {
                orc_sax(v.data(), r.data(), zeta_foll[k], 0., veo, 3);
                if (orc_norm2(v.data(), veo, 3) < prec) {
                    conv[k] = 1;
                    ++nconv;
                }
            }
        }
        zeta_prev = zeta;
        zeta      = zeta_foll;
        tmp1      = tmp2;
        if (nconv == nshift) {
            if (tmp2 != tmp2) {
                *iters = it;
                return ORC_SOLVER_STUCK;
            }
            if (tmp2 < prec) {
                *iters = it;
                return ORC_OK;
            }
        }
        ++it;
    }
    *iters = it;
    return ORC_SOLVER_DID_NOT_SOLVE;
}

}  // extern "C"
