// A batch of independent P3M systems (uneven N, mesh and order) evaluated
// with one launch per stage: every batched kernel spans all systems, so a
// batch of small systems fills the GPU where one small system cannot.  The
// sequence of launches can be captured once into a HIP graph and replayed.
//
// Memory: one arena for all meshes (charge / potential mesh, spectra, ik
// field meshes, sized for FP64 so the FP32 form fits too), one for the
// influence functions, Dn, self-force coefficients and twiddles; particle
// and cell sorts are global over the concatenated particle arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "batch_desc.hpp"
#include "k_internal.hpp"
#include "particle_bodies.hpp"

using namespace p3m;

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename V> int upload(V **dst, const std::vector<V> &src) {
    P3MK_TRY(p3mk_alloc((void **)dst, sizeof(V) * std::max<size_t>(src.size(), 1)));
    hipError_t e = hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "batch descriptor upload", (int)e);
    return P3MK_OK;
}

}  // namespace

struct p3mk_batch {
    BForm form;
    int nsys = 0;
    int ntot = 0;
    std::vector<p3mk_bsys_in> in;
    std::vector<BSys> hs;
    // device descriptors
    BSys *dsys = nullptr;
    int *psys = nullptr, *p0s = nullptr, *cell0s = nullptr;
    long *tile0s = nullptr, *blk_z = nullptr, *blk_y = nullptr, *blk_x = nullptr;
    long nblk_z = 0, nblk_y = 0, nblk_x = 0;
    // LDS of the FFT passes, of the interpolation (a tile's halo) and of
    // the row-parallel assignment (region + staging)
    size_t lds_z = 0, lds_y = 0, lds_x = 0, lds_assign = 0, lds_rows = 0;
    long ntiles = 0, max_items = 0, ncell = 0;
    int tile_side = 8; // requested assignment tile side (autotuned)
    int rs_grid = 0;   // persistent real-space grid (autotuned)
    bool time_pairs = false; // events around the half-shell kernel (eager calls only)
    bool rs_small = false, rs_cells = false; // systems with nearest-image / cell-grid real space
    int cus = 256;
    // arenas
    void *mesh_arena = nullptr;
    size_t mesh_region = 0; // bytes of the charge meshes (zeroed per call)
    void *aux_arena = nullptr;
    // particle sort
    int *key = nullptr, *count = nullptr, *start = nullptr, *cursor = nullptr, *perm = nullptr,
        *chunks = nullptr, *cstart = nullptr, *scan = nullptr;
    int4 *imap = nullptr; // work item -> (tile, chunk, system)
    double4 *sorted = nullptr;
    // cells
    CellArrays cells;
    // twiddles of each system (host copies kept for precision changes)
    std::vector<size_t> tw_off;
};

namespace {

void release_all(p3mk_batch *b) {
    // cursor lives in the count allocation
    void *ptrs[] = {b->dsys,  b->psys,   b->p0s,        b->cell0s,    b->tile0s, b->blk_z,
                    b->blk_y, b->blk_x,  b->mesh_arena, b->aux_arena, b->key,    b->count,
                    b->start, b->perm,   b->chunks,     b->cstart,    b->scan,   b->sorted,
                    b->imap};
    for (void *p : ptrs)
        p3mk_release(p);
    cells_release(b->cells);
}

// Build the tile-dependent descriptors (tile side T per system) and upload
// every descriptor array.
int build_descriptors(p3mk_batch *b) {
    const int ns = b->nsys;
    std::vector<long> tile0s(ns + 1), blk_z(ns + 1), blk_y(ns + 1), blk_x(ns + 1);
    long tiles = 0, items = 0, bz = 0, by = 0, bx = 0;
    size_t lds_a = 0, lds_r = 0, lz = 0, ly = 0, lx = 0;
    const int NM = b->form.interlaced ? 2 : 1;
    for (int s = 0; s < ns; s++) {
        BSys &S = b->hs[s];
        S.T = S.M < b->tile_side ? S.M : b->tile_side;
        S.nt = (S.M + S.T - 1) / S.T;
        S.tile0 = tiles;
        tile0s[s] = tiles;
        tiles += (long)S.nt * S.nt * S.nt;
        items += (S.n + part::kChunk - 1) / part::kChunk + (long)S.nt * S.nt * S.nt;
        const int R = S.T + S.cao - 1 + (b->form.interlaced ? 1 : 0);
        lds_a = std::max(lds_a, sizeof(double) * NM * (size_t)R * R * R);
        const part::RowRegion rg = part::row_region(S.T, S.cao, b->form.interlaced != 0);
        lds_r = std::max(lds_r, sizeof(double) * (size_t)rg.size +
                                    (size_t)(part::kThreads / 64) *
                                        part::row_stage_bytes_n(S.cao, b->form.fp32 ? 4 : 8));
        const int M = S.M, W = b->form.r2c ? M / 2 + 1 : M;
        const int l2 = bfft_lines_per_tile(b->form, M, kPassZ);
        const int lxx = bfft_lines_per_tile(b->form, M, kPassX);
        blk_z[s] = bz;
        bz += ((long)M * M + l2 - 1) / l2;
        blk_y[s] = by;
        by += ((long)M * W + l2 - 1) / l2;
        blk_x[s] = bx;
        bx += ((long)M * W + lxx - 1) / lxx;
        lz = std::max(lz, bfft_lds_bytes(b->form, M, kPassZ));
        ly = std::max(ly, bfft_lds_bytes(b->form, M, kPassY));
        lx = std::max(lx, bfft_lds_bytes(b->form, M, kPassX));
    }
    tile0s[ns] = tiles;
    blk_z[ns] = bz;
    blk_y[ns] = by;
    blk_x[ns] = bx;
    b->ntiles = tiles;
    b->max_items = items;
    b->nblk_z = bz;
    b->nblk_y = by;
    b->nblk_x = bx;
    b->lds_z = lz;
    b->lds_y = ly;
    b->lds_x = lx;
    b->lds_assign = lds_a;
    b->lds_rows = lds_r;
    if (lds_a > 64 * 1024 || lds_r > 64 * 1024)
        return p3mk_fail(P3MK_EINVAL, "assignment tile does not fit in LDS", 0);
    // (re)upload the descriptor arrays
    p3mk_release(b->dsys);
    p3mk_release(b->tile0s);
    p3mk_release(b->blk_z);
    p3mk_release(b->blk_y);
    p3mk_release(b->blk_x);
    b->dsys = nullptr;
    b->tile0s = b->blk_z = b->blk_y = b->blk_x = nullptr;
    P3MK_TRY(upload(&b->dsys, b->hs));
    P3MK_TRY(upload(&b->tile0s, tile0s));
    P3MK_TRY(upload(&b->blk_z, blk_z));
    P3MK_TRY(upload(&b->blk_y, blk_y));
    P3MK_TRY(upload(&b->blk_x, blk_x));
    // particle sort over all tiles
    void **sort_ptrs[] = {(void **)&b->count, (void **)&b->start, (void **)&b->chunks,
                          (void **)&b->cstart, (void **)&b->scan, (void **)&b->imap};
    for (void **p : sort_ptrs) {
        p3mk_release(*p);
        *p = nullptr;
    }
    b->cursor = nullptr;
    const size_t nt1 = (size_t)tiles + 1;
    P3MK_TRY(alloc_count_cursor(&b->count, &b->cursor, nt1));
    P3MK_TRY(p3mk_alloc((void **)&b->start, sizeof(int) * nt1));
    P3MK_TRY(p3mk_alloc((void **)&b->chunks, sizeof(int) * nt1));
    P3MK_TRY(p3mk_alloc((void **)&b->cstart, sizeof(int) * nt1));
    P3MK_TRY(p3mk_alloc((void **)&b->imap, sizeof(int4) * (size_t)(b->max_items > 0 ? b->max_items : 1)));
    P3MK_TRY(p3mk_alloc((void **)&b->scan, sizeof(int) * p3mk_scan_scratch_ints(tiles)));
    return P3MK_OK;
}

}  // namespace

extern "C" void p3mk_batch_destroy(p3mk_batch *b) {
    if (!b)
        return;
    release_all(b);
    delete b;
}

extern "C" int p3mk_batch_create(p3mk_batch **out, int r2c, int interlaced, int ad, int fp32,
                                 int nsys, const p3mk_bsys_in *sys) {
    *out = nullptr;
    if (nsys < 1 || sys == nullptr)
        return p3mk_fail(P3MK_EINVAL, "a batch needs at least one system", 0);
    P3MK_TRY(p3mk_device_check());
    p3mk_batch *b = new (std::nothrow) p3mk_batch();
    if (!b)
        return p3mk_fail(P3MK_ENOMEM, "batch allocation", 0);
    b->form = {r2c != 0, interlaced != 0, fp32 != 0, ad != 0};
    b->nsys = nsys;
    b->in.assign(sys, sys + nsys);
    b->hs.resize(nsys);
    {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        b->cus = cus > 0 ? cus : 256;
        b->rs_grid = b->cus * 32;
    }
    // sizes and offsets
    p3mk_geom g64 = {0, 0, r2c, interlaced, 0, 1.0};
    size_t mesh_off = 0, aux_off = 0;
    std::vector<size_t> moff(nsys), spec_off(nsys), f_off(nsys), g_off(nsys), dn_off(nsys),
        sb_off(nsys);
    b->tw_off.resize(nsys);
    long ntot = 0, ncell = 0;
    std::vector<int> p0s(nsys + 1), cell0s(nsys + 1);
    for (int s = 0; s < nsys; s++) {
        const p3mk_bsys_in &I = sys[s];
        fft::Radices rp;
        if (I.n < 0 || I.M < 2 || I.M % 2 != 0 || !fft::make_radices(I.M, rp) || I.cao < 1 ||
            I.cao > 7 || I.M < I.cao || !(I.L > 0.0) || I.rcut > 0.5 * I.L) {
            delete b;
            return p3mk_fail(P3MK_EINVAL, "batch system with an invalid mesh, order, size or box", 0);
        }
        BSys &S = b->hs[s];
        S.M = I.M;
        S.cao = I.cao;
        S.L = I.L;
        S.Hi = (double)I.M / I.L;
        S.h = I.L / I.M;
        S.p0 = (int)ntot;
        S.n = I.n;
        S.rp = rp;
        S.prefac = I.prefac;
        S.rcut = I.rcut;
        S.alpha = I.alpha;
        p0s[s] = (int)ntot;
        ntot += I.n;
        g64.M = I.M;
        moff[s] = mesh_off;
        mesh_off += align256(p3mk_mesh_bytes(&g64));
        int nc = 1, reach = 1;
        if (I.rcut > 0.0)
            cell_geometry(I.n, I.L, I.rcut, nc, reach);
        S.nc = nc;
        S.reach = reach;
        S.small = (2 * reach + 1 > nc) ? 1 : 0;
        if (I.rcut > 0.0) {
            b->rs_small = b->rs_small || S.small;
            b->rs_cells = b->rs_cells || !S.small;
        }
        S.cell0 = (int)ncell;
        cell0s[s] = (int)ncell;
        ncell += (long)nc * nc * nc;
    }
    p0s[nsys] = (int)ntot;
    cell0s[nsys] = (int)ncell;
    if (ntot > kCellMaxParticles || ncell > 0x7fffffffL / 16) {
        delete b;
        return p3mk_fail(P3MK_EINVAL, "batch too large", 0);
    }
    b->ntot = (int)ntot;
    b->ncell = ncell;
    b->mesh_region = mesh_off;
    // spectra and field meshes after the charge meshes
    for (int s = 0; s < nsys; s++) {
        g64.M = b->hs[s].M;
        spec_off[s] = mesh_off;
        if (r2c)
            mesh_off += (ad ? 1 : 3) * align256(p3mk_spec_bytes(&g64));
        f_off[s] = mesh_off;
        if (!ad)
            mesh_off += (r2c ? 2 : 3) * align256(p3mk_mesh_bytes(&g64));
        const size_t M = (size_t)b->hs[s].M;
        g_off[s] = aux_off;
        aux_off += align256(sizeof(double) * M * M * M);
        dn_off[s] = aux_off;
        aux_off += align256(sizeof(double) * M);
        sb_off[s] = aux_off;
        aux_off += align256(sizeof(double) * 81);
        b->tw_off[s] = aux_off;
        aux_off += align256(sizeof(double2) * M) + align256(sizeof(float2) * M);
    }
    int rc = p3mk_alloc(&b->mesh_arena, mesh_off);
    if (rc == P3MK_OK) rc = p3mk_alloc(&b->aux_arena, aux_off);
    if (rc != P3MK_OK) {
        p3mk_batch_destroy(b);
        return rc;
    }
    char *ma = (char *)b->mesh_arena, *aa = (char *)b->aux_arena;
    for (int s = 0; s < nsys; s++) {
        BSys &S = b->hs[s];
        g64.M = S.M;
        const size_t mb = align256(p3mk_mesh_bytes(&g64)), sbb = align256(p3mk_spec_bytes(&g64));
        for (int k = 0; k < kNumBuf; k++)
            S.buf[k] = nullptr;
        S.buf[kBufMesh] = ma + moff[s];
        if (r2c)
            for (int k = 0; k < (ad ? 1 : 3); k++)
                S.buf[kBufSpec0 + k] = ma + spec_off[s] + k * sbb;
        if (!ad) {
            if (r2c) {
                S.buf[kBufF0] = S.buf[kBufMesh];
                S.buf[kBufF0 + 1] = ma + f_off[s];
                S.buf[kBufF0 + 2] = ma + f_off[s] + mb;
            } else {
                for (int k = 0; k < 3; k++)
                    S.buf[kBufF0 + k] = ma + f_off[s] + k * mb;
            }
        }
        S.G = (const double *)(aa + g_off[s]);
        S.Dn = (const double *)(aa + dn_off[s]);
        S.selfb = nullptr;
        // twiddles exp(-2 pi i k / M), from long double
        std::vector<double2> h64(S.M);
        std::vector<float2> h32(S.M);
        for (int k = 0; k < S.M; k++) {
            const long double a = -2.0L * 3.141592653589793238462643383279502884L * k / S.M;
            h64[k] = make_double2((double)cosl(a), (double)sinl(a));
            h32[k] = make_float2((float)cosl(a), (float)sinl(a));
        }
        char *tw = aa + b->tw_off[s];
        hipError_t e = hipMemcpy(tw, h64.data(), sizeof(double2) * S.M, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(tw + align256(sizeof(double2) * S.M), h32.data(), sizeof(float2) * S.M,
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            p3mk_batch_destroy(b);
            return p3mk_fail(P3MK_EDEVICE, "twiddle upload", (int)e);
        }
        S.tw = fp32 ? (const void *)(tw + align256(sizeof(double2) * S.M)) : (const void *)tw;
    }
    // per-particle system ids and prefix arrays
    std::vector<int> psys((size_t)std::max<long>(ntot, 1));
    for (int s = 0; s < nsys; s++)
        std::fill(psys.begin() + p0s[s], psys.begin() + p0s[s + 1], s);
    const size_t nn = (size_t)std::max<long>(ntot, 1);
    rc = upload(&b->psys, psys);
    if (rc == P3MK_OK) rc = upload(&b->p0s, p0s);
    if (rc == P3MK_OK) rc = upload(&b->cell0s, cell0s);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&b->key, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&b->perm, sizeof(int) * nn);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&b->sorted, sizeof(double4) * nn);
    if (rc == P3MK_OK) rc = cells_alloc(b->cells, ntot, ncell);
    if (rc == P3MK_OK) rc = build_descriptors(b);
    if (rc != P3MK_OK) {
        p3mk_batch_destroy(b);
        return rc;
    }
    *out = b;
    return P3MK_OK;
}

extern "C" int p3mk_batch_system_arrays(p3mk_batch *b, int s, double **G, double **Dn,
                                        double **selfb) {
    if (s < 0 || s >= b->nsys)
        return p3mk_fail(P3MK_EINVAL, "system index out of range", 0);
    BSys &S = b->hs[s];
    *G = (double *)S.G;
    *Dn = (double *)S.Dn;
    // the self-force coefficients live 81 doubles after Dn's aligned block
    *selfb = (double *)((char *)S.Dn + align256(sizeof(double) * S.M));
    return P3MK_OK;
}

extern "C" int p3mk_batch_configure(p3mk_batch *b, int fp32, int use_self_force, int tile_side,
                                    int rs_grid) {
    if (tile_side != 4 && tile_side != 8)
        return p3mk_fail(P3MK_EINVAL, "assignment tile side must be 4 or 8", 0);
    b->form.fp32 = fp32 != 0;
    b->tile_side = tile_side;
    if (rs_grid > 0)
        b->rs_grid = rs_grid;
    char *aa = (char *)b->aux_arena;
    for (int s = 0; s < b->nsys; s++) {
        BSys &S = b->hs[s];
        char *tw = aa + b->tw_off[s];
        S.tw = fp32 ? (const void *)(tw + align256(sizeof(double2) * S.M)) : (const void *)tw;
        S.selfb = use_self_force
                      ? (const double *)((const char *)S.Dn + align256(sizeof(double) * S.M))
                      : nullptr;
    }
    return build_descriptors(b);
}

extern "C" int p3mk_batch_set_alpha(p3mk_batch *b, int s, double alpha) {
    if (s < 0 || s >= b->nsys)
        return p3mk_fail(P3MK_EINVAL, "system index out of range", 0);
    b->hs[s].alpha = alpha;
    hipError_t e = hipMemcpy(&b->dsys[s], &b->hs[s], sizeof(BSys), hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return p3mk_fail(P3MK_EDEVICE, "batch descriptor upload", (int)e);
    return P3MK_OK;
}

extern "C" int p3mk_batch_info(const p3mk_batch *b, int *ntot, long *ntiles, long *ncell,
                               int *tile_side, int *rs_grid) {
    if (ntot) *ntot = b->ntot;
    if (ntiles) *ntiles = b->ntiles;
    if (ncell) *ncell = b->ncell;
    if (tile_side) *tile_side = b->tile_side;
    if (rs_grid) *rs_grid = b->rs_grid;
    return P3MK_OK;
}

// k-space force of every system: fk += (3 blocks of ntot), ftot -= self force
// (methods with the correction)
extern "C" int p3mk_batch_kspace(p3mk_batch *b, const double *x, const double *y, const double *z,
                                 const double *q, double *fk, double *ftot, p3mk_queue *qu) {
    const BForm &f = b->form;
    const int ns = b->nsys;
    P3MK_TRY(p3mk_fill_zero(b->mesh_arena, b->mesh_region, qu));
    P3MK_TRY(bpart_sort(f, b->dsys, b->psys, b->tile0s, ns, b->ntiles, b->ntot, x, y, z, q, b->key,
                        b->count, b->start, b->cursor, b->perm, b->chunks, b->cstart, b->imap,
                        b->scan, b->sorted, qu));
    P3MK_TRY(bpart_assign(f, b->dsys, b->ntiles, b->max_items, b->lds_rows, b->start, b->cstart,
                          b->imap, b->sorted, qu));
    // forward transform along z and y
    int spec = kBufMesh;
    if (f.r2c) {
        spec = kBufSpec0;
        P3MK_TRY(bfft_r2c_z(f, b->dsys, b->blk_z, ns, b->nblk_z, b->lds_z, kBufMesh, kBufSpec0, qu));
    } else {
        P3MK_TRY(bfft_pass(f, b->dsys, b->blk_z, ns, b->nblk_z, b->lds_z, kPassZ, kBufMesh, 0, qu));
    }
    P3MK_TRY(bfft_pass(f, b->dsys, b->blk_y, ns, b->nblk_y, b->lds_y, kPassY, spec, 0, qu));
    if (!f.ad) {
        const int outs = f.r2c ? kBufSpec0 : kBufF0;
        P3MK_TRY(bfft_convolve_x(f, b->dsys, b->blk_x, ns, b->nblk_x, b->lds_x, 3, spec, outs, qu));
        for (int k = 0; k < 3; k++) {
            P3MK_TRY(bfft_pass(f, b->dsys, b->blk_y, ns, b->nblk_y, b->lds_y, kPassY, outs + k, 1, qu));
            if (f.r2c)
                P3MK_TRY(bfft_c2r_z(f, b->dsys, b->blk_z, ns, b->nblk_z, b->lds_z, outs + k,
                                    kBufF0 + k, qu));
            else
                P3MK_TRY(bfft_pass(f, b->dsys, b->blk_z, ns, b->nblk_z, b->lds_z, kPassZ, outs + k,
                                   1, qu));
        }
    } else {
        P3MK_TRY(bfft_convolve_x(f, b->dsys, b->blk_x, ns, b->nblk_x, b->lds_x, 1, spec, spec, qu));
        P3MK_TRY(bfft_pass(f, b->dsys, b->blk_y, ns, b->nblk_y, b->lds_y, kPassY, spec, 1, qu));
        if (f.r2c)
            P3MK_TRY(bfft_c2r_z(f, b->dsys, b->blk_z, ns, b->nblk_z, b->lds_z, spec, kBufMesh, qu));
        else
            P3MK_TRY(bfft_pass(f, b->dsys, b->blk_z, ns, b->nblk_z, b->lds_z, kPassZ, kBufMesh, 1, qu));
    }
    return bpart_interp(f, b->dsys, b->ntiles, b->max_items, b->lds_assign, b->start, b->cstart,
                        b->imap, b->sorted, b->perm, b->ntot, fk, ftot, qu);
}

// real-space force of every system with rcut > 0: fr += (3 blocks of ntot);
// energy[s] += system s's real-space energy (energy may be NULL)
extern "C" int p3mk_batch_realspace(p3mk_batch *b, const double *x, const double *y,
                                    const double *z, const double *q, double *fr, double *energy,
                                    p3mk_queue *qu) {
    return brs_forces(b->form, b->dsys, b->psys, b->cell0s, b->nsys, b->ncell, b->ntot, x, y, z, q,
                      b->cells, b->rs_grid, b->rs_small, b->rs_cells, b->time_pairs, fr, energy, qu);
}

extern "C" int p3mk_batch_time_pairs(p3mk_batch *b, int on) {
    b->time_pairs = on != 0;
    return P3MK_OK;
}
