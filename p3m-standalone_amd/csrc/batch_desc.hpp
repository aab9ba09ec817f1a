// Descriptors of a batch of independent systems and the batched launchers of
// the FFT, particle and real-space kernels (C++ linkage, internal to the
// library).  Every batched launch covers all systems: a block (or a particle
// slot, or a work item) finds its system by binary search in a prefix array,
// then runs the same per-tile / per-particle / per-item body as the
// single-system kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fft_tiles.hpp"
#include "k_internal.hpp"

namespace p3m {

// buffer slots of a system's mesh arena
enum BufSlot { kBufMesh = 0, kBufSpec0 = 1, kBufF0 = 4, kNumBuf = 7 };

struct BSys {
    // mesh geometry
    int M, cao, T, nt;  // mesh, order, tile side, tiles per side
    double L, Hi, h;
    long tile0;         // first global tile index
    int p0, n;          // particle slots [p0, p0 + n) in the concatenated arrays
    fft::Radices rp;
    void *buf[kNumBuf]; // mesh, spec0..2, F0..2 (element type of the batch)
    const double *G, *Dn, *selfb;
    const void *tw;     // twiddles exp(-2 pi i k / M), batch precision
    double prefac;      // back-interpolation prefactor (ad: includes nothing else)
    // real space
    int cell0, nc, reach, small;
    double rcut, alpha;
};

// forms shared by all systems of a batch
struct BForm {
    int r2c, interlaced, fp32, ad;
};

// ---- launchers (defined next to the single-system kernels) ----
// FFT: blk = block prefix over systems for the given pass
int bfft_pass(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks,
              size_t lds, int axis,
              int slot, int inv, p3mk_queue *q);
int bfft_r2c_z(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks,
              size_t lds, int in_slot,
               int out_slot, p3mk_queue *q);
int bfft_c2r_z(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks,
              size_t lds, int in_slot,
               int out_slot, p3mk_queue *q);
int bfft_convolve_x(const BForm &f, const BSys *sys, const long *blk, int nsys, long nblocks,
              size_t lds,
                    int nout, int in_slot, int out_slot, p3mk_queue *q);
// lines per tile and LDS buffers of each batched pass kind (host side)
int bfft_lines_per_tile(const BForm &f, int M, int kind);
size_t bfft_lds_bytes(const BForm &f, int M, int kind);
enum { kPassZ = 0, kPassY = 1, kPassX = 2 };

// particles: sort by global tile, assign, back-interpolate
// imap: work item -> (global tile, chunk, system); max_items entries
int bpart_sort(const BForm &f, const BSys *sys, const int *psys, const long *tile0s, int nsys,
               long ntiles, int ntot, const double *x, const double *y, const double *z,
               const double *q, int *key, int *count, int *start, int *cursor, int *perm,
               int *chunks, int *cstart, int4 *imap, int *scan_scratch, double4 *sorted,
               p3mk_queue *qu);
int bpart_assign(const BForm &f, const BSys *sys, long ntiles, long max_items, size_t lds,
                 const int *start, const int *cstart, const int4 *imap, const double4 *sorted,
                 p3mk_queue *qu);
int bpart_interp(const BForm &f, const BSys *sys, long ntiles, long max_items, size_t lds,
                 const int *start, const int *cstart, const int4 *imap, const double4 *sorted,
                 const int *perm, int ntot, double *fk, double *ftot, p3mk_queue *qu);

// real space: cell-list arrays (octant-sorted cells, work items of 64
// particles with bounding boxes) and one persistent launch over all items
// Most particles a cell list holds: the real-space kernels address the
// sorted FP64 positions (and 64 padding entries) with 32-bit byte offsets.
constexpr long kCellMaxParticles = (1L << 27) - 64;

// Counts and cursors of a counting sort over m keys in one allocation
// (cursor = count + m rounded up to a multiple of 4), cleared by one fill.
inline int alloc_count_cursor(int **count, int **cursor, size_t m) {
    const size_t m4 = (m + 3) & ~(size_t)3;
    *cursor = nullptr;
    const int rc = p3mk_alloc((void **)count, sizeof(int) * 2 * m4);
    if (rc == P3MK_OK)
        *cursor = *count + m4;
    return rc;
}

inline int zero_count_cursor(int *count, int *cursor, size_t m, p3mk_queue *qu) {
    const size_t m4 = (m + 3) & ~(size_t)3;
    if (cursor == count + m4)
        return p3mk_fill_zero(count, sizeof(int) * 2 * m4, qu);
    P3MK_TRY(p3mk_fill_zero(count, sizeof(int) * m, qu));
    return p3mk_fill_zero(cursor, sizeof(int) * m, qu);
}

// Below this mean number of particles per cell of side rcut / 2, the cell
// list takes cells of side rcut instead.  In a sparse system a cell of side
// rcut / 2 holds one or two particles, and each work item's walk over its 63
// neighbour cells then outweighs its pairs.  Measured with 200k uniform
// particles at rcut 3, FP64 forces: 2 per cell take 1.24 ms with small cells
// against 0.39 with large ones, 8 per cell 0.81 against 0.55, 16 per cell
// 0.85 against 0.85 and 32 per cell 1.24 against 1.38.  The mean misjudges
// clustered systems: a Gaussian of 1M particles in a box of side 100 has 8
// per cell on average but about 90 in the typical particle's cell, and takes
// 11.0 ms with small cells against 12.0 with large ones.
constexpr double kSmallCellMinOccupancy = 6.0;
constexpr int kMaxCellsPerSide = 512;

// Cell grid of a real-space cell list for n particles: nc cells per side of
// side >= rcut / reach, with a margin so that rounding of x nc / L never puts
// a pair inside the cutoff more than `reach` cells apart.  Boxes of fewer than
// 2 reach + 1 cells per side are left to the nearest-image search.
inline void cell_geometry(long n, double L, double rcut, int &nc, int &reach) {
    const double margin = 1.0 - 1e-12;
    const int nc2 = (int)std::min<double>(kMaxCellsPerSide, std::floor(L * 2.0 / rcut * margin));
    const int nc1 = (int)std::min<double>(kMaxCellsPerSide, std::floor(L / rcut * margin));
    const bool sparse = nc1 >= 3 && (double)n < kSmallCellMinOccupancy * nc2 * (double)nc2 * nc2;
    if (nc2 < 5 || sparse) {
        reach = 1;
        nc = nc1;
    } else {
        reach = 2;
        nc = nc2;
    }
    if (nc < 1)
        nc = 1;
}

struct CellArrays {
    int *key = nullptr;                               // fine key per particle, then its sorted position
    int *fcount = nullptr, *fstart = nullptr; // 8 ncell + 1
    int *start = nullptr, *items = nullptr, *istart = nullptr;    // ncell + 1
    int *perm = nullptr, *scan = nullptr, *counter = nullptr;
    int *rank = nullptr; // per particle: its rank within its fine key (from the count's atomic)
    double4 *sorted = nullptr;
    // relative to the cell centre in FP32: (-2 x, -2 y, -2 z, q); the sorted
    // arrays carry 64 padding entries past the last particle
    float4 *sorted32 = nullptr;
    // relative to the cell centre in fixed point (29 fractional bits per
    // side for reach 2, 30 for reach 1): (x, y, z, q as FP32 bits)
    int4 *sortedq = nullptr;
    double4 *gsorted = nullptr; // adjoint cotangent in sorted order (made on first use)
    double *fsorted = nullptr;  // half-shell forces in sorted order, 3 n
    float4 *bbox = nullptr;     // 2 per work item
    long max_items = 0;
};
int cells_alloc(CellArrays &a, long n, long ncell);
void cells_release(CellArrays &a);
int brs_forces(const BForm &f, const BSys *sys, const int *psys, const int *cell0s, int nsys,
               long ncell, int ntot, const double *x, const double *y, const double *z,
               const double *q, CellArrays &a, int grid, bool any_small, bool any_cells,
               bool time_pairs, double *fr, double *energy, p3mk_queue *qu);

}  // namespace p3m
