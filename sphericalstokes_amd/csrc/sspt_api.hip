// sspt_api.hip -- the extern "C" boundary of libsspt.so (declared in
// include/sspt.h): context and buffers, metric tables, layout-permuting
// upload/download, per-kernel entry points and the PT loops.
#include "sspt_internal.h"
#include "sspt_launch.h"
#include "sspt_tables.h"
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace sspt;

namespace {

int fail(ss_ctx *c, int code, const char *msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_fail(ss_ctx *c, hipError_t e, const char *where)
{
    if (c) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
        c->err = buf;
    }
    return (int)e;
}

#define HIP_TRY(c, call)                                          \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call);   \
    } while (0)

bool is_shear(int f) { return f == SS_TAU_RTH || f == SS_TAU_RPH || f == SS_TAU_THPH; }

// Reference shape of a field and its box offset (1 for shear stresses).
void dims_of(const Box &b, int f, int d[3], int *o1)
{
    int nr = b.nr, nt = b.nt, np = b.np;
    *o1 = 0;
    switch (f) {
    case SS_VR:
    case SS_VR_ERR: d[0] = nr + 1; d[1] = nt; d[2] = np; break;
    case SS_VTH:
    case SS_VTH_ERR: d[0] = nr; d[1] = nt + 1; d[2] = np; break;
    case SS_VPH:
    case SS_VPH_ERR: d[0] = nr; d[1] = nt; d[2] = np + 1; break;
    case SS_TAU_RTH: d[0] = nr - 1; d[1] = nt - 1; d[2] = np - 2; *o1 = 1; break;
    case SS_TAU_RPH: d[0] = nr - 1; d[1] = nt - 2; d[2] = np - 1; *o1 = 1; break;
    case SS_TAU_THPH: d[0] = nr - 2; d[1] = nt - 1; d[2] = np - 1; *o1 = 1; break;
    default: d[0] = nr; d[1] = nt; d[2] = np; break;
    }
}

int sync(ss_ctx *c, const char *where)
{
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, where);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, where);
    return SS_OK;
}

int check_launch(ss_ctx *c, const char *where)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, where);
    return SS_OK;
}

// Fields that the fused loop rewrites every iteration and whose neighbours
// are read (the PT state).  The other fields are read-only inputs (rho,
// eta_ini), outputs of the last iteration (etaSM, DT_R, tII, V_err) or
// updated strictly point-wise in place (eta_PL) and are shared by both sets.
bool is_pingpong(int f)
{
    switch (f) {
    case SS_P: case SS_TAU_RR: case SS_TAU_THTH: case SS_TAU_PHPH:
    case SS_TAU_RTH: case SS_TAU_RPH: case SS_TAU_THPH:
    case SS_VR: case SS_VTH: case SS_VPH: case SS_ETA:
        return true;
    default:
        return false;
    }
}

// Allocates a buffer set; with `share`, non-ping-pong fields alias it.
// The current set allocates its fields; the ping-pong set shares the
// non-state fields.  With pair_state, a state field's two buffers are the
// halves of one allocation: the current set allocates both, the ping-pong
// set takes the upper half.
int alloc_set(ss_ctx *c, FieldSet &s, const FieldSet *share)
{
    for (int f = 0; f < SS_NFIELDS; ++f) {
        if (share && !is_pingpong(f)) {
            s.f[f] = share->f[f];
            continue;
        }
        if (c->pair_state && is_pingpong(f)) {
            if (share) {
                s.f[f] = c->pair_base[f] + c->box_elems;
                continue;
            }
            void *p = nullptr;
            if (hipMalloc(&p, 2 * c->box_elems * sizeof(double)) != hipSuccess) {
                c->err = "hipMalloc of a field box pair failed";
                return SS_ERR_NOMEM;
            }
            c->pair_base[f] = (double *)p;
            s.f[f] = (double *)p;
            hipError_t e = hipMemsetAsync(p, 0, 2 * c->box_elems * sizeof(double), c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "hipMemsetAsync");
            continue;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, c->box_elems * sizeof(double));
        if (e != hipSuccess) {
            c->err = "hipMalloc of a field box failed";
            return SS_ERR_NOMEM;
        }
        s.f[f] = (double *)p;
        e = hipMemsetAsync(p, 0, c->box_elems * sizeof(double), c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "hipMemsetAsync");
    }
    return SS_OK;
}

void free_set(ss_ctx *c, FieldSet &s, const FieldSet *share)
{
    for (int f = 0; f < SS_NFIELDS; ++f) {
        if (s.f[f] && !c->pair_base[f] && !(share && s.f[f] == share->f[f])) (void)hipFree(s.f[f]);
        s.f[f] = nullptr;
    }
}

// Device pointers of the table block, in the order build_tables fills it.
void bind_tables(ss_ctx *c)
{
    const int n1 = c->box.nr + 1, n2 = c->box.nt + 1;
    const double *p = c->tab_mem;
    auto take = [&](size_t n) { const double *q = p; p += n; return q; };
    Tables &t = c->tab;
    t.Rc = take(n1); t.invRc = take(n1); t.Rf = take(n1); t.invRf = take(n1);
    t.Re = take(n1); t.invRe = take(n1); t.Rc4 = take(n1); t.invRc4 = take(n1);
    t.Rc2 = take(n1); t.invRc2 = take(n1);
    t.cotc = take(n2); t.cotf4 = take(n2); t.cotc2 = take(n2); t.cotf2 = take(n2);
    t.rs_c = take((size_t)n1 * n2); t.rs_e = take((size_t)n1 * n2); t.rs_c4 = take((size_t)n1 * n2);
    t.rs_f = take((size_t)n1 * n2); t.rs_y = take((size_t)n1 * n2); t.rs_z = take((size_t)n1 * n2);
    t.th = take(n2); t.ph = take(c->box.np + 1);
}

} // namespace

extern "C" {

int ss_create(int nr, int nth, int nph, int device, ss_ctx **out)
{
    if (!out) return SS_ERR_ARG;
    *out = nullptr;
    if (nr < 3 || nth < 3 || nph < 3 || nr > 4096 || nth > 4096 || nph > 4096) return SS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return (int)hipErrorNoDevice;
    if (device < 0 || device >= ndev) return SS_ERR_ARG;
    ss_ctx *c = new ss_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete c; return (int)e; }
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
            c->n_cu = ncu;
    }
    Box &b = c->box;
    b.nr = nr; b.nt = nth; b.np = nph;
    b.pitch = ((nph + 1) + 63) / 64 * 64;
    b.sj = b.pitch;
    b.si = (long long)(nth + 1) * b.pitch;
    c->box_elems = (size_t)(nr + 1) * (size_t)b.si;
    // the fused kernels address a field with 32-bit byte offsets
    if (c->box_elems * sizeof(double) > 0xFFFFFFFFull) { delete c; return SS_ERR_ARG; }
    if (const char *ev = std::getenv("SSPT_PAIRED")) c->pair_state = std::atoi(ev) != 0;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return (int)e; }
    int rc = alloc_set(c, c->cur, nullptr);
    if (rc != SS_OK) { ss_destroy(c); return rc; }
    c->staging_elems = (size_t)(nr + 1) * (nth + 1) * (nph + 1);
    if (hipMalloc(&c->staging, c->staging_elems * sizeof(double)) != hipSuccess) { ss_destroy(c); return SS_ERR_NOMEM; }
    if (hipMalloc(&c->red, 16 * sizeof(unsigned long long)) != hipSuccess) { ss_destroy(c); return SS_ERR_NOMEM; }
    if (hipMalloc(&c->tbp, 2 * (size_t)nr * nth * sizeof(double)) != hipSuccess) { ss_destroy(c); return SS_ERR_NOMEM; }
    if (hipHostMalloc(&c->red_host, 16 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        ss_destroy(c);
        return SS_ERR_NOMEM;
    }
    if (hipMalloc(&c->tab_mem, tables_elems(b) * sizeof(double)) != hipSuccess) { ss_destroy(c); return SS_ERR_NOMEM; }
    bind_tables(c);
    if (hipMalloc(&c->ftab_mem, fused_tables_elems(b) * sizeof(double)) != hipSuccess) { ss_destroy(c); return SS_ERR_NOMEM; }
    c->ftr = c->ftab_mem;
    c->ftt = c->ftr + (size_t)FT_NTR * (nr + 1);
    c->ft2 = c->ftt + (size_t)FT_NTT * (nth + 1);
    {
        const size_t n32 = fused_tables_elems(b) + (nth + 1) + (nph + 1);
        if (hipMalloc(&c->ftab32_mem, n32 * sizeof(float)) != hipSuccess) { ss_destroy(c); return SS_ERR_NOMEM; }
        c->ftr32 = c->ftab32_mem;
        c->ftt32 = c->ftr32 + (size_t)FT_NTR * (nr + 1);
        c->ft232 = c->ftt32 + (size_t)FT_NTT * (nth + 1);
        c->th32 = c->ft232 + (size_t)FT_NT2 * (nr + 1) * (nth + 1);
        c->ph32 = c->th32 + (nth + 1);
    }
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    rc = sync(c, "ss_create");
    if (rc != SS_OK) { ss_destroy(c); return rc; }
    *out = c;
    return SS_OK;
}

int ss_destroy(ss_ctx *c)
{
    if (!c) return SS_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->alt_ready) free_set(c, c->alt, &c->cur);
    free_set(c, c->cur, nullptr);
    for (double *&p : c->pair_base)
        if (p) { (void)hipFree(p); p = nullptr; }
    if (c->staging) (void)hipFree(c->staging);
    if (c->red) (void)hipFree(c->red);
    if (c->tbp) (void)hipFree(c->tbp);
    if (c->red_host) (void)hipHostFree(c->red_host);
    if (c->tab_mem) (void)hipFree(c->tab_mem);
    if (c->ftab_mem) (void)hipFree(c->ftab_mem);
    if (c->ftab32_mem) (void)hipFree(c->ftab32_mem);
    for (double *p : c->tau_scr)
        if (p) (void)hipFree(p);
    for (int q = 0; q < SS_F32_NBOX; ++q) // the correction pairs are freed through their lower halves
        if (c->f32[q] && !(q < 20 && c->pair_base32[q % 10])) (void)hipFree(c->f32[q]);
    for (float *p : c->pair_base32)
        if (p) (void)hipFree(p);
    for (double *p : c->adj)
        if (p) (void)hipFree(p);
    for (float *p : c->adj32)
        if (p) (void)hipFree(p);
    if (c->adj_part) (void)hipFree(c->adj_part);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SS_OK;
}

const char *ss_last_error(const ss_ctx *c)
{
    if (!c) return "null context";
    return c->err.c_str();
}

int ss_field_dims(const ss_ctx *c, int field, int dims[3])
{
    if (!c || !dims || field < 0 || field >= SS_NFIELDS) return SS_ERR_ARG;
    int o1;
    dims_of(c->box, field, dims, &o1);
    return SS_OK;
}

int ss_set_grid(ss_ctx *c, const double *R1, const double *th1, const double *ph1)
{
    if (!c) return SS_ERR_ARG;
    if (!R1 || !th1 || !ph1) return fail(c, SS_ERR_ARG, "ss_set_grid: null coordinate array");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<double> h;
    build_tables(c->box, R1, th1, ph1, h);
    HIP_TRY(c, hipMemcpyAsync(c->tab_mem, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> hf;
    build_fused_tables(c->box, h, hf);
    HIP_TRY(c, hipMemcpyAsync(c->ftab_mem, hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // FP32 copies for the correction sweeps (tables, then th and ph)
    std::vector<float> h32(hf.begin(), hf.end());
    const double *thp = h.data() + (h.size() - (c->box.nt + 1) - (c->box.np + 1));
    h32.insert(h32.end(), thp, thp + (c->box.nt + 1) + (c->box.np + 1));
    HIP_TRY(c, hipMemcpyAsync(c->ftab32_mem, h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    c->th_lo = th1[0];
    c->th_hi = th1[c->box.nt - 1];
    int rc = sync(c, "ss_set_grid");
    if (rc == SS_OK) c->grid_set = true;
    return rc;
}

int ss_upload(ss_ctx *c, int field, const double *host, int layout)
{
    if (!c) return SS_ERR_ARG;
    if (!host || field < 0 || field >= SS_NFIELDS || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C))
        return fail(c, SS_ERR_ARG, "ss_upload: bad argument");
    HIP_TRY(c, hipSetDevice(c->device));
    int d[3], o1;
    dims_of(c->box, field, d, &o1);
    const size_t n = (size_t)d[0] * d[1] * d[2];
    HIP_TRY(c, hipMemcpyAsync(c->staging, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    launch_to_box(c, c->cur.f[field], c->staging, d[0], d[1], d[2], o1, layout);
    return sync(c, "ss_upload");
}

int ss_download(ss_ctx *c, int field, double *host, int layout)
{
    if (!c) return SS_ERR_ARG;
    if (!host || field < 0 || field >= SS_NFIELDS || (layout != SS_LAYOUT_JULIA && layout != SS_LAYOUT_C))
        return fail(c, SS_ERR_ARG, "ss_download: bad argument");
    HIP_TRY(c, hipSetDevice(c->device));
    int d[3], o1;
    dims_of(c->box, field, d, &o1);
    const size_t n = (size_t)d[0] * d[1] * d[2];
    launch_from_box(c, c->staging, c->cur.f[field], d[0], d[1], d[2], o1, layout);
    HIP_TRY(c, hipMemcpyAsync(host, c->staging, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return sync(c, "ss_download");
}

int ss_fill(ss_ctx *c, int field, double value)
{
    if (!c) return SS_ERR_ARG;
    if (field < 0 || field >= SS_NFIELDS) return fail(c, SS_ERR_ARG, "ss_fill: bad field");
    HIP_TRY(c, hipSetDevice(c->device));
    int d[3], o1;
    dims_of(c->box, field, d, &o1);
    launch_fill(c, c->cur.f[field], d[0], d[1], d[2], o1, value);
    return sync(c, "ss_fill");
}

int ss_copy_field(ss_ctx *c, int dst, int src)
{
    if (!c) return SS_ERR_ARG;
    if (dst < 0 || dst >= SS_NFIELDS || src < 0 || src >= SS_NFIELDS) return fail(c, SS_ERR_ARG, "ss_copy_field: bad field");
    int d1[3], d2[3], o1, o2;
    dims_of(c->box, dst, d1, &o1);
    dims_of(c->box, src, d2, &o2);
    if (d1[0] != d2[0] || d1[1] != d2[1] || d1[2] != d2[2] || o1 != o2)
        return fail(c, SS_ERR_ARG, "ss_copy_field: shapes differ");
    HIP_TRY(c, hipSetDevice(c->device));
    launch_copy_field(c, c->cur.f[dst], c->cur.f[src], d1[0], d1[1], d1[2], o1);
    return sync(c, "ss_copy_field");
}

#define NEED_GRID(c)                                                                     \
    do {                                                                                 \
        if (!(c)) return SS_ERR_ARG;                                                     \
        if (!(c)->grid_set) return fail((c), SS_ERR_STATE, "ss_set_grid was not called"); \
        HIP_TRY((c), hipSetDevice((c)->device));                                         \
    } while (0)

int ss_initialize_inclusion(ss_ctx *c, int field, double r, double radius, double in)
{
    NEED_GRID(c);
    int d[3], o1;
    if (field < 0 || field >= SS_NFIELDS) return fail(c, SS_ERR_ARG, "bad field");
    dims_of(c->box, field, d, &o1);
    if (d[0] != c->box.nr || d[1] != c->box.nt || d[2] != c->box.np)
        return fail(c, SS_ERR_ARG, "initialize_inclusion needs a cell-centre field");
    launch_init_inclusion(c, c->cur.f[field], r, radius, in);
    return sync(c, "ss_initialize_inclusion");
}

int ss_initialize_velocity(ss_ctx *c, double fact, double r)
{
    NEED_GRID(c);
    launch_init_velocity(c, c->cur.f[SS_VPH], fact, r);
    return sync(c, "ss_initialize_velocity");
}

int ss_equal3(ss_ctx *c)
{
    if (!c) return SS_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    launch_equal3(c, c->cur);
    return sync(c, "ss_equal3");
}

int ss_maxloc(ss_ctx *c)
{
    if (!c) return SS_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    launch_maxloc(c, c->cur.f[SS_ETASM], c->cur.f[SS_ETA]);
    return sync(c, "ss_maxloc");
}

int ss_spread_etasm(ss_ctx *c, int m)
{
    if (!c) return SS_ERR_ARG;
    if (m < 1 || m > SS_SPREAD_MAX) return fail(c, SS_ERR_ARG, "ss_spread_etasm: m must be in [1, 8]");
    HIP_TRY(c, hipSetDevice(c->device));
    launch_etasm_spread(c, c->cur.f[SS_ETASM], c->cur.f[SS_DT_R], m);
    return sync(c, "ss_spread_etasm");
}

int ss_copy_bc(ss_ctx *c, int field)
{
    if (!c) return SS_ERR_ARG;
    if (field < 0 || field >= SS_NFIELDS || is_shear(field)) return fail(c, SS_ERR_ARG, "ss_copy_bc: bad field");
    HIP_TRY(c, hipSetDevice(c->device));
    int d[3], o1;
    dims_of(c->box, field, d, &o1);
    launch_copy_bc(c, c->cur.f[field], d[0], d[1], d[2]);
    return sync(c, "ss_copy_bc");
}

int ss_compute_P(ss_ctx *c, double _dr, double _dth, double _dph, double dmp, double vpdt, double lr)
{
    NEED_GRID(c);
    launch_compute_P(c, c->cur, _dr, _dth, _dph, dmp, vpdt, lr);
    return sync(c, "ss_compute_P");
}

int ss_compute_TII(ss_ctx *c)
{
    if (!c) return SS_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    launch_compute_TII(c, c->cur);
    return sync(c, "ss_compute_TII");
}

int ss_power_law(ss_ctx *c, double tau_C, double n_exp)
{
    if (!c) return SS_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    launch_power_law(c, c->cur, tau_C, n_exp, 1.0, 0.0, INFINITY);
    return sync(c, "ss_power_law");
}

int ss_power_law_ex(ss_ctx *c, double tau_C, double n_exp, double relax, double eta_min, double eta_max)
{
    if (!c) return SS_ERR_ARG;
    if (!(relax > 0.0 && relax <= 1.0)) return fail(c, SS_ERR_ARG, "relax must be in (0,1]");
    HIP_TRY(c, hipSetDevice(c->device));
    launch_power_law(c, c->cur, tau_C, n_exp, relax, eta_min, eta_max);
    return sync(c, "ss_power_law_ex");
}

int ss_compute_V(ss_ctx *c, double _dr, double _dth, double _dph, double g)
{
    NEED_GRID(c);
    launch_compute_V(c, c->cur, _dr, _dth, _dph, g);
    return sync(c, "ss_compute_V");
}

int ss_bc_R(ss_ctx *c)
{
    if (!c) return SS_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    launch_bc_R(c, c->cur);
    return sync(c, "ss_bc_R");
}

int ss_bc_th(ss_ctx *c, int mode, double r, double sh)
{
    NEED_GRID(c);
    if (mode < SS_BC_SHEAR || mode > SS_BC_PURE_SHEAR) return fail(c, SS_ERR_ARG, "bad bc_mode");
    launch_bc_th(c, c->cur, mode, r, sh);
    return sync(c, "ss_bc_th");
}

int ss_bc_ph(ss_ctx *c, int mode, double r, double sh)
{
    NEED_GRID(c);
    if (mode < SS_BC_SHEAR || mode > SS_BC_PURE_SHEAR) return fail(c, SS_ERR_ARG, "bad bc_mode");
    launch_bc_ph(c, c->cur, mode, r, sh);
    return sync(c, "ss_bc_ph");
}

} // extern "C"

namespace {

double jl_max_h(double x, double y)
{
    bool pick_y = (y > x) || (std::signbit(y) < std::signbit(x));
    return pick_y ? (std::isnan(x) ? x : y) : (std::isnan(y) ? y : x);
}

// Reads the six reduction slots and forms err as SphericalStokes_diapir.jl:429-432.
int finish_err(ss_ctx *c, int drop_nan_vph, double comp[3], double *err)
{
    hipError_t e = hipMemcpyAsync(c->red_host, c->red, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "reduction readback");
    int rc = sync(c, "check_err");
    if (rc != SS_OK) return rc;
    double me[3], mv[3];
    for (int a = 0; a < 3; ++a) {
        memcpy(&me[a], &c->red_host[a], sizeof(double));
        memcpy(&mv[a], &c->red_host[3 + a], sizeof(double));
        comp[a] = me[a] / mv[a];
    }
    if (drop_nan_vph && std::isnan(comp[2]))
        *err = jl_max_h(comp[0], comp[1]);
    else
        *err = jl_max_h(jl_max_h(comp[0], comp[1]), comp[2]);
    return SS_OK;
}

int validate_params(ss_ctx *c, const ss_params *p)
{
    if (!p) return fail(c, SS_ERR_ARG, "null params");
    if (p->bc_mode < SS_BC_SHEAR || p->bc_mode > SS_BC_PURE_SHEAR) return fail(c, SS_ERR_ARG, "bad bc_mode");
    if (!(p->relax > 0.0 && p->relax <= 1.0)) return fail(c, SS_ERR_ARG, "relax must be in (0,1]");
    if (p->etasm_spread < 0 || p->etasm_spread > SS_SPREAD_MAX)
        return fail(c, SS_ERR_ARG, "etasm_spread must be in [0, 8]");
    return SS_OK;
}

// One PT iteration through the per-kernel launchers, reference order
// (SphericalStokes_diapir.jl:413-425); equal3! only where its copy is read.
void unfused_step(ss_ctx *c, const ss_params *p, bool do_equal3)
{
    const FieldSet &s = c->cur;
    const Box &b = c->box;
    if (do_equal3) launch_equal3(c, s);
    launch_maxloc(c, s.f[SS_ETASM], s.f[SS_ETA]);
    launch_copy_bc(c, s.f[SS_ETASM], b.nr, b.nt, b.np);
    // spread etaSM (build option); DT_R is the temporary, compute_P! rewrites all of it next
    if (p->etasm_spread > 0) launch_etasm_spread(c, s.f[SS_ETASM], s.f[SS_DT_R], p->etasm_spread);
    launch_compute_P(c, s, p->_dr, p->_dth, p->_dph, p->dmp, p->vpdt, p->lr);
    launch_compute_TII(c, s);
    launch_copy_bc(c, s.f[SS_TII], b.nr, b.nt, b.np);
    launch_power_law(c, s, p->tau_C, p->n_exp, p->relax, p->eta_min, p->eta_max);
    launch_compute_V(c, s, p->_dr, p->_dth, p->_dph, p->g);
    launch_bc_R(c, s);
    launch_bc_th(c, s, p->bc_mode, p->r, p->sh);
    launch_bc_ph(c, s, p->bc_mode, p->r, p->sh);
}

using clk = std::chrono::steady_clock;

} // namespace

extern "C" {

int ss_check_err(ss_ctx *c, int drop_nan_vph, double err_comp[3], double *err)
{
    if (!c) return SS_ERR_ARG;
    if (!err_comp || !err) return fail(c, SS_ERR_ARG, "ss_check_err: null output");
    HIP_TRY(c, hipSetDevice(c->device));
    const FieldSet &s = c->cur;
    launch_check_err(c, s, s.f[SS_VR_ERR], s.f[SS_VTH_ERR], s.f[SS_VPH_ERR], 1);
    return finish_err(c, drop_nan_vph, err_comp, err);
}

int ss_pt_solve_unfused(ss_ctx *c, const ss_params *p, int maxiter, int nout, double eps, ss_result *out)
{
    NEED_GRID(c);
    int rc = validate_params(c, p);
    if (rc != SS_OK) return rc;
    if (maxiter < 0 || nout < 0 || !out) return fail(c, SS_ERR_ARG, "bad loop arguments");
    memset(out, 0, sizeof *out);
    out->err = NAN;
    clk::time_point t0 = clk::now();
    int ittot = 0;
    bool nonfinite = false;
    for (int iter = 1; iter <= maxiter; ++iter) {
        if (iter == 11) {
            if ((rc = sync(c, "pt loop")) != SS_OK) return rc;
            t0 = clk::now();
        }
        const bool check = nout > 0 && iter % nout == 0;
        unfused_step(c, p, check);
        if ((rc = check_launch(c, "pt loop")) != SS_OK) return rc;
        if (check) {
            const FieldSet &s = c->cur;
            launch_check_err(c, s, s.f[SS_VR_ERR], s.f[SS_VTH_ERR], s.f[SS_VPH_ERR], 1);
            if ((rc = finish_err(c, p->err_drop_nan_vph, out->err_comp, &out->err)) != SS_OK) return rc;
            if (std::isnan(out->err)) { nonfinite = true; ittot = iter; break; }
            if (out->err < eps && iter > 20) { out->iter_end = iter; break; }
        }
        ittot = iter;
    }
    if ((rc = sync(c, "pt loop")) != SS_OK) return rc;
    out->seconds = std::chrono::duration<double>(clk::now() - t0).count();
    out->ittot = ittot;
    if (nonfinite) return fail(c, SS_ERR_NONFINITE, "NaN in the convergence error");
    if (nout > 0 && out->iter_end == 0 && maxiter > 0) return SS_NOT_CONVERGED;
    return SS_OK;
}

} // extern "C"

// ---------------------------------------------------------------- fused loop
namespace {

int ensure_alt(ss_ctx *c)
{
    if (c->alt_ready) return SS_OK;
    int rc = alloc_set(c, c->alt, &c->cur);
    if (rc != SS_OK) { free_set(c, c->alt, &c->cur); return rc; }
    c->alt_ready = true;
    return sync(c, "alloc ping-pong");
}

} // namespace

extern "C" {

int ss_pt_iterate(ss_ctx *c, const ss_params *p, int niter)
{
    NEED_GRID(c);
    int rc = validate_params(c, p);
    if (rc != SS_OK) return rc;
    if (niter < 0) return fail(c, SS_ERR_ARG, "niter < 0");
    if (niter == 0) return SS_OK;
    if ((rc = ensure_alt(c)) != SS_OK) return rc;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    for (int it = 0; it < niter; ++it) {
        launch_fused_iteration(c, c->cur, c->alt, *p, it == niter - 1, false);
        std::swap(c->cur, c->alt);
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    if ((rc = sync(c, "ss_pt_iterate")) != SS_OK) return rc;
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = (double)ms / niter;
    return SS_OK;
}

int ss_pt_solve(ss_ctx *c, const ss_params *p, int maxiter, int nout, double eps, ss_result *out)
{
    NEED_GRID(c);
    int rc = validate_params(c, p);
    if (rc != SS_OK) return rc;
    if (maxiter < 0 || nout < 0 || !out) return fail(c, SS_ERR_ARG, "bad loop arguments");
    if ((rc = ensure_alt(c)) != SS_OK) return rc;
    memset(out, 0, sizeof *out);
    out->err = NAN;
    clk::time_point t0 = clk::now();
    int ittot = 0;
    bool nonfinite = false;
    for (int iter = 1; iter <= maxiter; ++iter) {
        if (iter == 11) {
            if ((rc = sync(c, "pt loop")) != SS_OK) return rc;
            t0 = clk::now();
        }
        const bool check = nout > 0 && iter % nout == 0;
        // the state after the loop must include etaSM/DT_R/tII of the last
        // iteration; it is the last one unless the exit test fires earlier,
        // so a check iteration also materialises them.
        const bool last = (iter == maxiter) || check;
        launch_fused_iteration(c, c->cur, c->alt, *p, last, check);
        std::swap(c->cur, c->alt);
        if ((rc = check_launch(c, "pt loop")) != SS_OK) return rc;
        if (check) {
            if ((rc = finish_err(c, p->err_drop_nan_vph, out->err_comp, &out->err)) != SS_OK) return rc;
            if (std::isnan(out->err)) { nonfinite = true; ittot = iter; break; }
            if (out->err < eps && iter > 20) { out->iter_end = iter; break; }
        }
        ittot = iter;
    }
    if ((rc = sync(c, "pt loop")) != SS_OK) return rc;
    out->seconds = std::chrono::duration<double>(clk::now() - t0).count();
    out->ittot = ittot;
    if (nonfinite) return fail(c, SS_ERR_NONFINITE, "NaN in the convergence error");
    if (nout > 0 && out->iter_end == 0 && maxiter > 0) return SS_NOT_CONVERGED;
    return SS_OK;
}

int ss_pt_iterate_timed(ss_ctx *c, const ss_params *p, int niter, double *main_ms, double *iter_ms)
{
    NEED_GRID(c);
    int rc = validate_params(c, p);
    if (rc != SS_OK) return rc;
    if (niter < 1 || !main_ms || !iter_ms) return fail(c, SS_ERR_ARG, "ss_pt_iterate_timed: bad argument");
    if ((rc = ensure_alt(c)) != SS_OK) return rc;
    while ((int)c->ev_pool.size() < 2 * niter) {
        hipEvent_t e;
        HIP_TRY(c, hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    for (int it = 0; it < niter; ++it) {
        launch_fused_iteration(c, c->cur, c->alt, *p, it == niter - 1, false, c->ev_pool[2 * it], c->ev_pool[2 * it + 1]);
        std::swap(c->cur, c->alt);
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    if ((rc = sync(c, "ss_pt_iterate_timed")) != SS_OK) return rc;
    double sum = 0.0;
    for (int it = 0; it < niter; ++it) {
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev_pool[2 * it], c->ev_pool[2 * it + 1]));
        sum += ms;
    }
    float tot = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&tot, c->ev0, c->ev1));
    *main_ms = sum / niter;
    *iter_ms = (double)tot / niter;
    c->last_kernel_ms = *iter_ms;
    return SS_OK;
}

int ss_last_kernel_ms(const ss_ctx *c, double *ms)
{
    if (!c || !ms) return SS_ERR_ARG;
    *ms = c->last_kernel_ms;
    return SS_OK;
}

} // extern "C"
