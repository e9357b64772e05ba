// Cell keys, stable cell sort and particle permutation.
//
// The key is the reference's flat index (ix*Ny + iy)*Nz + iz with
// ix = clip(floor((x + W/2)/dx), 0, Nx-1) (particle_initialization.py:89-102,
// particle_tile_communication.py:62-79), evaluated in FP64 from the FP32
// position so it matches an FP64 evaluation bit for bit.  Species are kept in
// their own segments and inactive slots sort to the end of their segment.  The
// permutation is rocPRIM's LSD radix sort of (key, slot) pairs, which is
// stable, so it equals np.argsort(key, kind="stable") (:159).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "pic_common.h"

extern "C" int pic_fill_f32(float* ptr, int64_t n, float value, void* stream);

namespace {

constexpr int kBlock = 256;

__global__ void cell_keys_kernel(PicDesc d, PicSpecies sp, PicParticles in, uint32_t* keys, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = species_of(i, sp.slots, sp.count);
    int64_t ncell = (int64_t)d.n[0] * d.n[1] * d.n[2];
    int64_t key = ncell;
    if (in.active[i]) {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            double t = floor(((double)in.x[a][i] + 0.5 * d.wind[a]) / d.d[a]);
            t = fmin(fmax(t, 0.0), (double)(d.n[a] - 1));      // NaN -> 0 via fmax
            c[a] = (int)t;
        }
        key = ((int64_t)c[0] * d.n[1] + c[1]) * d.n[2] + c[2];
    }
    keys[i] = (uint32_t)(s * (ncell + 1) + key);
}

// Tile-major cell key: cubic tiles of T cells per axis, tiles in x-major
// order, cells inside a tile in x-major order; inactive slots get the key
// after the last tile.  Same cell formula as cell_keys_kernel.
__host__ __device__ inline void tile_counts(const PicDesc& d, int T, int nt[3]) {
    for (int a = 0; a < 3; ++a) nt[a] = (d.n[a] + T - 1) / T;
}

// Per-axis constants of the tile key: W/2 and 1/d in FP64, the tile edge as a
// shift when it is a power of two (else -1).
struct TileKeyConst {
    double half_w[3], inv_d[3];
    int nt[3], shift;
};

__device__ inline uint32_t tile_key(const PicDesc& d, int T, const TileKeyConst& kc, int s, bool act,
                                    const float x[3]) {
    const int64_t vol = (int64_t)T * T * T;
    const int64_t span = (int64_t)kc.nt[0] * kc.nt[1] * kc.nt[2] * vol;
    int64_t key = span;
    if (act) {
        int c[3], tq[3], tr[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double t = floor(((double)x[a] + kc.half_w[a]) * kc.inv_d[a]);
            t = fmin(fmax(t, 0.0), (double)(d.n[a] - 1));
            c[a] = (int)t;
            tq[a] = kc.shift >= 0 ? c[a] >> kc.shift : c[a] / T;
            tr[a] = kc.shift >= 0 ? c[a] & (T - 1) : c[a] % T;
        }
        const int64_t tile = ((int64_t)tq[0] * kc.nt[1] + tq[1]) * kc.nt[2] + tq[2];
        const int64_t local = ((int64_t)tr[0] * T + tr[1]) * T + tr[2];
        key = tile * vol + local;
    }
    return (uint32_t)(s * (span + 1) + key);
}

__global__ void tile_keys_kernel(PicDesc d, PicSpecies sp, PicParticles in, int T, TileKeyConst kc, uint32_t* keys,
                                 int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = (uint32_t)i;     // n < 2^31 is checked at launch
    const float x[3] = {in.x[0][u], in.x[1][u], in.x[2][u]};
    keys[u] = tile_key(d, T, kc, species_of(i, sp.slots, sp.count), in.active[u] != 0, x);
}

// ---- counting re-bin by tile key -------------------------------------------
// A block takes kRebinChunk consecutive slots of the current order and
// aggregates their keys in an LDS hash table (nearly sorted input has a few
// dozen distinct keys per chunk, plus the particles that changed cell), so the
// global atomics are one per distinct key per chunk.  Equal keys in
// consecutive lanes of a 16-lane row are first counted with DPP moves
// (row_run_position), so only the last lane of such a run touches the table.
// Pass 1 counts the keys and stores them; after the scan, pass 2 reserves each
// key's positions for the chunk (one atomic per key, one LDS atomic per run)
// and writes the inverse permutation
// perm[dst] = src; the particles are then gathered through perm (reads of
// nearby slots, coalesced writes).  The order inside one key follows the
// atomics and is not reproducible; the kernels do not depend on it.

constexpr int kRebinItems = 4;
constexpr int kRebinChunk = kBlock * kRebinItems;
// one slot per item: a chunk of nearly sorted particles has a few dozen keys, and
// the table cannot overflow (a random order fills it, at the cost of long probes)
constexpr int kHashSize = kRebinChunk;
static_assert((kHashSize & (kHashSize - 1)) == 0, "the hash table size is a power of two");
constexpr uint32_t kEmpty = 0xffffffffu;

__device__ inline int hash_slot(uint32_t key) {
    return (int)((key * 2654435761u) >> (32 - __builtin_ctz(kHashSize))) & (kHashSize - 1);
}

// slot of key in the table (inserting it), by linear probing
__device__ inline int hash_insert(uint32_t* tkey, uint32_t key) {
    int h = hash_slot(key);
    for (;;) {
        const uint32_t old = atomicCAS(&tkey[h], kEmpty, key);
        if (old == kEmpty || old == key) return h;
        h = (h + 1) & (kHashSize - 1);
    }
}

__global__ void rebin_count_kernel(PicDesc d, PicSpecies sp, PicParticles in, int T, TileKeyConst kc, float ahead,
                                   uint32_t* __restrict__ count, uint32_t* __restrict__ keys, int64_t n) {
    __shared__ uint32_t tkey[kHashSize];
    __shared__ uint32_t tcnt[kHashSize];
    for (int e = threadIdx.x; e < kHashSize; e += kBlock) {
        tkey[e] = kEmpty;
        tcnt[e] = 0u;
    }
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kRebinChunk;
    // every item's loads first, so they share one memory round trip
    float X[kRebinItems][3], U[kRebinItems][3];
    bool A[kRebinItems];
#pragma unroll
    for (int k = 0; k < kRebinItems; ++k) {
        const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
        A[k] = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) X[k][a] = U[k][a] = 0.0f;
        if (i < n) {
            A[k] = in.active[i] != 0;
#pragma unroll
            for (int a = 0; a < 3; ++a) X[k][a] = in.x[a][i];
            if (ahead != 0.0f) {        // uniform
#pragma unroll
                for (int a = 0; a < 3; ++a) U[k][a] = in.u[a][i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kRebinItems; ++k) {
        const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
        const bool in_range = i < n;
        uint32_t key = kEmpty;          // past the end: a run of its own, never inserted
        if (in_range) {
            const int s = species_of(i, sp.slots, sp.count);
            float x[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) x[a] = sp.update_x[s][a] ? fmaf(ahead, U[k][a], X[k][a]) : X[k][a];
            key = tile_key(d, T, kc, s, A[k], x);
            keys[i] = key;
        }
        int p, q;
        row_run_position((int)key, p, q);
        if (in_range && q == 1) atomicAdd(&tcnt[hash_insert(tkey, key)], (uint32_t)p);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kHashSize; e += kBlock)
        if (tcnt[e]) atomicAdd(&count[tkey[e]], tcnt[e]);
}

__global__ void rebin_perm_kernel(const uint32_t* __restrict__ keys, uint32_t* __restrict__ pos,
                                  int32_t* __restrict__ perm, int64_t n) {
    __shared__ uint32_t tkey[kHashSize];
    __shared__ uint32_t tcnt[kHashSize];
    for (int e = threadIdx.x; e < kHashSize; e += kBlock) {
        tkey[e] = kEmpty;
        tcnt[e] = 0u;
    }
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kRebinChunk;
    int slot[kRebinItems];
    uint32_t rank[kRebinItems];
    const int lane = (int)(threadIdx.x & 63u);
    for (int k = 0; k < kRebinItems; ++k) {
        const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
        const bool in_range = i < n;
        const uint32_t key = in_range ? keys[i] : kEmpty;
        int p, q;
        row_run_position((int)key, p, q);
        // the run's last lane inserts the key and reserves the run's p ranks;
        // the other lanes of the run take its slot and first rank
        int sl = -1;
        uint32_t base = 0u;
        if (in_range && q == 1) {
            sl = hash_insert(tkey, key);
            base = atomicAdd(&tcnt[sl], (uint32_t)p);
        }
        const int src = lane + q - 1;
        sl = __shfl(sl, src);
        base = (uint32_t)__shfl((int)base, src);
        slot[k] = in_range ? sl : -1;
        rank[k] = base + (uint32_t)(p - 1);
    }
    // the chunk's base for each of its keys: the occupied table slots are listed
    // first, so the returning global atomics go out in one round
    __shared__ int nlist;
    __shared__ uint16_t list[kHashSize];
    if (threadIdx.x == 0) nlist = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < kHashSize; e += kBlock) {
        const bool used = tcnt[e] != 0u;
        const unsigned long long m = __ballot(used);
        int first = 0;
        if (lane == 0 && m) first = atomicAdd(&nlist, __popcll(m));
        first = __shfl(first, 0);
        if (used) list[first + __popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)e;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nlist; q += kBlock) {
        const int e = list[q];
        tcnt[e] = atomicAdd(&pos[tkey[e]], tcnt[e]);
    }
    __syncthreads();
    for (int k = 0; k < kRebinItems; ++k) {
        if (slot[k] < 0) continue;
        const int64_t i = b0 + (int64_t)k * kBlock + threadIdx.x;
        perm[tcnt[slot[k]] + rank[k]] = (int32_t)i;
    }
}

// Exclusive scan of n uint32 counts in place: per-block sums, a scan of the
// sums, then each block scans its items from its offset.
constexpr int kScanItems = 16;
constexpr int kScanChunk = kBlock * kScanItems;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const uint32_t add = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const uint32_t incl = sh[t];
    total = sh[kBlock - 1];
    __syncthreads();
    return incl - v;
}

__global__ void scan_sums_kernel(const uint32_t* __restrict__ v, int64_t n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[kBlock];
    const int64_t b0 = (int64_t)blockIdx.x * kScanChunk;
    uint32_t acc = 0u;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t e = b0 + (int64_t)k * kBlock + threadIdx.x;
        if (e < n) acc += v[e];
    }
    uint32_t total;
    block_exclusive_scan(acc, sh, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// one block: exclusive scan of the block sums in place
__global__ void scan_block_sums_kernel(uint32_t* __restrict__ sums, int64_t nb) {
    __shared__ uint32_t sh[kBlock];
    uint32_t carry = 0u;
    for (int64_t base = 0; base < nb; base += kBlock) {
        const int64_t e = base + threadIdx.x;
        const uint32_t v = e < nb ? sums[e] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, total);
        if (e < nb) sums[e] = carry + ex;
        carry += total;
    }
}

__global__ void scan_apply_kernel(uint32_t* __restrict__ v, int64_t n, const uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[kBlock];
    const int64_t b0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    uint32_t loc[kScanItems];
    uint32_t acc = 0u;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t e = b0 + k;
        loc[k] = e < n ? v[e] : 0u;
        acc += loc[k];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(acc, sh, total) + sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t e = b0 + k;
        if (e < n) v[e] = off;
        off += loc[k];
    }
}

// starts[b] = first sorted slot of (species, tile) b = the exclusive prefix
// at that pair's first key; starts[last] = n
__global__ void rebin_starts_kernel(const uint32_t* __restrict__ pos, int64_t span, int64_t vol, int32_t tiles,
                                    int32_t species, int64_t n, int32_t* __restrict__ starts) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)species * (tiles + 1);
    if (b > nb) return;
    if (b == nb) {
        starts[b] = (int32_t)n;
        return;
    }
    const int64_t s = b / (tiles + 1), t = b % (tiles + 1);
    starts[b] = (int32_t)pos[s * (span + 1) + t * vol];
}

// starts[b] = first sorted position whose key >= the first key of tile b,
// b = s * (tiles + 1) + t (t = tiles: the inactive tail); starts[last] = n.
__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys, int64_t n, int64_t span, int64_t vol,
                                   int32_t tiles, int32_t species, int32_t* starts) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)species * (tiles + 1);
    if (b > nb) return;
    if (b == nb) {
        starts[b] = (int32_t)n;
        return;
    }
    const int64_t s = b / (tiles + 1), t = b % (tiles + 1);
    const uint64_t target = (uint64_t)(s * (span + 1) + t * vol);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    starts[b] = (int32_t)lo;
}

// permute_kernel for four consecutive destinations per thread: one 16-byte
// load of perm, the 28 gathered loads issued together, 16-byte stores per plane
// (planes 16-byte aligned, activity bytes 4-byte aligned); a last group of
// fewer than four takes the per-slot path.
__global__ void permute4_kernel(PicParticles in, PicParticles out, const int32_t* __restrict__ perm, int64_t n) {
    // XCD-contiguous destination ranges: a particle that moved to a nearby
    // cell is read from slots the same XCD has just streamed through its L2
    const int64_t blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int64_t i0 = 4 * (blk * blockDim.x + threadIdx.x);
    if (i0 >= n) return;
    if (i0 + 4 <= n) {
        const int4 p4 = *reinterpret_cast<const int4*>(perm + i0);
        int64_t src[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (src[j] < 0 || src[j] >= n) src[j] = i0 + j;     // defensive: never read outside the planes
        float v[6][4];
        uint32_t act = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                v[a][j] = in.x[a][src[j]];
                v[3 + a][j] = in.u[a][src[j]];
            }
            act |= (uint32_t)in.active[src[j]] << (8 * j);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            store_vec<4>(out.x[a] + i0, v[a]);
            store_vec<4>(out.u[a] + i0, v[3 + a]);
        }
        *reinterpret_cast<uint32_t*>(out.active + i0) = act;
        return;
    }
    for (int64_t i = i0; i < n; ++i) {
        int64_t src = perm[i];
        if (src < 0 || src >= n) src = i;
        for (int a = 0; a < 3; ++a) {
            out.x[a][i] = in.x[a][src];
            out.u[a][i] = in.u[a][src];
        }
        out.active[i] = in.active[src];
    }
}

__global__ void permute_kernel(PicParticles in, PicParticles out, const int32_t* __restrict__ perm, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t src = perm[i];
    if (src < 0 || src >= n) src = i;   // defensive: never read outside the planes
    for (int a = 0; a < 3; ++a) {
        out.x[a][i] = in.x[a][src];
        out.u[a][i] = in.u[a][src];
    }
    out.active[i] = in.active[src];
}

size_t sort_storage(int64_t n, int bits) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                              rocprim::counting_iterator<int32_t>(0), (int32_t*)nullptr, (size_t)n, 0u,
                              (unsigned)bits, (hipStream_t)0);
    return bytes;
}

}  // namespace

extern "C" {

int64_t pic_tile_count(const PicDesc* desc, int32_t tile) {
    if (!desc || tile < 1) return -1;
    int nt[3];
    tile_counts(*desc, tile, nt);
    return (int64_t)nt[0] * nt[1] * nt[2];
}

int pic_tile_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t tile, int32_t* keys,
                  void* stream) {
    PIC_REQUIRE(desc && keys && tile >= 1 && particles_ok(sp, in));
    const int64_t span = pic_tile_count(desc, tile) * tile * tile * tile;
    PIC_REQUIRE((int64_t)sp->count * (span + 1) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    if (n == 0) return 0;
    TileKeyConst kc;
    tile_counts(*desc, tile, kc.nt);
    kc.shift = -1;
    for (int b = 0; b < 31; ++b)
        if (tile == (1 << b)) kc.shift = b;
    for (int a = 0; a < 3; ++a) {
        kc.half_w[a] = 0.5 * desc->wind[a];
        kc.inv_d[a] = 1.0 / desc->d[a];
    }
    hipLaunchKernelGGL(tile_keys_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       (int)tile, kc, reinterpret_cast<uint32_t*>(keys), n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_tile_ranges(const PicDesc* desc, const PicSpecies* sp, int32_t tile, const int32_t* keys_sorted,
                    int32_t* starts, void* stream) {
    PIC_REQUIRE(desc && sp && starts && tile >= 1 && sp->count >= 0 && sp->count <= PIC_MAX_SPECIES);
    PIC_REQUIRE(sp->slots >= 0 && (keys_sorted || (int64_t)sp->count * sp->slots == 0));
    const int64_t tiles = pic_tile_count(desc, tile);
    const int64_t vol = (int64_t)tile * tile * tile;
    const int64_t n = (int64_t)sp->count * sp->slots;
    PIC_REQUIRE(n < ((int64_t)1 << 31) && (int64_t)sp->count * (tiles * vol + 1) < ((int64_t)1 << 31));
    const int64_t nb = (int64_t)sp->count * (tiles + 1) + 1;
    hipLaunchKernelGGL(tile_ranges_kernel, dim3(grid_1d(nb, kBlock)), dim3(kBlock), 0, as_stream(stream),
                       reinterpret_cast<const uint32_t*>(keys_sorted), n, tiles * vol, vol, (int32_t)tiles,
                       (int32_t)sp->count, starts);
    PIC_CHECK_LAUNCH();
    return 0;
}

static bool tile_key_const(const PicDesc* desc, int32_t tile, TileKeyConst& kc) {
    if (!desc || tile < 1) return false;
    tile_counts(*desc, tile, kc.nt);
    kc.shift = -1;
    for (int b = 0; b < 31; ++b)
        if (tile == (1 << b)) kc.shift = b;
    for (int a = 0; a < 3; ++a) {
        kc.half_w[a] = 0.5 * desc->wind[a];
        kc.inv_d[a] = 1.0 / desc->d[a];
    }
    return true;
}

static int64_t rebin_bins(const PicDesc* desc, const PicSpecies* sp, int32_t tile) {
    const int64_t tiles = pic_tile_count(desc, tile);
    if (tiles < 0 || !sp) return -1;
    return (int64_t)sp->count * (tiles * tile * tile * tile + 1);
}

int64_t pic_tile_rebin_workspace_bytes(const PicDesc* desc, const PicSpecies* sp, int32_t tile) {
    const int64_t bins = rebin_bins(desc, sp, tile);
    if (bins < 0) return -1;
    const int64_t blocks = (bins + kScanChunk - 1) / kScanChunk;
    const int64_t n = (int64_t)sp->count * sp->slots;
    return ((bins * 4 + 255) / 256) * 256 + ((blocks * 4 + 255) / 256) * 256 + n * 4 + 256;
}

int pic_tile_rebin(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t tile,
                   double lookahead, PicParticles* out, int32_t* starts, int32_t* perm, void* workspace,
                   int64_t workspace_bytes, void* stream) {
    TileKeyConst kc;
    PIC_REQUIRE(tile_key_const(desc, tile, kc) && particles_ok(sp, in) && starts);
    const int64_t n = (int64_t)sp->count * sp->slots;
    const int64_t bins = rebin_bins(desc, sp, tile);
    PIC_REQUIRE(n < ((int64_t)1 << 31) && bins < ((int64_t)1 << 31));
    PIC_REQUIRE(workspace && workspace_bytes >= pic_tile_rebin_workspace_bytes(desc, sp, tile));
    const int64_t tiles = pic_tile_count(desc, tile);
    const int64_t vol = (int64_t)tile * tile * tile;
    hipStream_t s = as_stream(stream);
    const int64_t nb = (int64_t)sp->count * (tiles + 1) + 1;
    if (n > 0) PIC_REQUIRE(perm);
    if (n > 0 && out) {
        for (int a = 0; a < 3; ++a)
            PIC_REQUIRE(out->x[a] && out->u[a] && out->x[a] != in->x[a] && out->u[a] != in->u[a]);
        PIC_REQUIRE(out->active && out->active != in->active);
    }
    char* w = static_cast<char*>(workspace);
    const int64_t blocks = (bins + kScanChunk - 1) / kScanChunk;
    uint32_t* pos = reinterpret_cast<uint32_t*>(w);
    uint32_t* sums = reinterpret_cast<uint32_t*>(w + ((bins * 4 + 255) / 256) * 256);
    uint32_t* keys = reinterpret_cast<uint32_t*>(w + ((bins * 4 + 255) / 256) * 256 + ((blocks * 4 + 255) / 256) * 256);
    int rc = pic_fill_f32(reinterpret_cast<float*>(pos), bins, 0.0f, stream);    // zero bits
    if (rc) return rc;
    const int64_t chunks = (n + kRebinChunk - 1) / kRebinChunk;
    if (n > 0) {
        PIC_REQUIRE(std::isfinite(lookahead));
        hipLaunchKernelGGL(rebin_count_kernel, dim3((unsigned)chunks), dim3(kBlock), 0, s, *desc, *sp, *in, (int)tile,
                           kc, (float)lookahead, pos, keys, n);
        PIC_CHECK_LAUNCH();
    }
    if (blocks > 0) {       // no species: no bins, only the closing entry of starts
        hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, pos, bins, sums);
        hipLaunchKernelGGL(scan_block_sums_kernel, dim3(1), dim3(kBlock), 0, s, sums, blocks);
        hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, pos, bins, sums);
    }
    hipLaunchKernelGGL(rebin_starts_kernel, dim3(grid_1d(nb, kBlock)), dim3(kBlock), 0, s, pos, tiles * vol, vol,
                       (int32_t)tiles, (int32_t)sp->count, n, starts);
    PIC_CHECK_LAUNCH();
    if (n > 0) {
        hipLaunchKernelGGL(rebin_perm_kernel, dim3((unsigned)chunks), dim3(kBlock), 0, s, keys, pos, perm, n);
        PIC_CHECK_LAUNCH();
    }
    if (n > 0 && out) {
        bool vec = (((uintptr_t)perm) & 15u) == 0 && (((uintptr_t)out->active) & 3u) == 0;
        for (int a = 0; a < 3; ++a) vec = vec && ((((uintptr_t)out->x[a]) | ((uintptr_t)out->u[a])) & 15u) == 0;
        if (vec)
            hipLaunchKernelGGL(permute4_kernel, dim3(grid_1d((n + 3) / 4, kBlock)), dim3(kBlock), 0, s, *in, *out,
                               perm, n);
        else
            hipLaunchKernelGGL(permute_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, s, *in, *out, perm, n);
        PIC_CHECK_LAUNCH();
    }
    return 0;
}

int pic_cell_keys(const PicDesc* desc, const PicSpecies* sp, const PicParticles* in, int32_t* keys, void* stream) {
    PIC_REQUIRE(desc && keys && particles_ok(sp, in));
    int64_t ncell = (int64_t)desc->n[0] * desc->n[1] * desc->n[2];
    PIC_REQUIRE((int64_t)sp->count * (ncell + 1) < ((int64_t)1 << 31));
    int64_t n = (int64_t)sp->count * sp->slots;
    if (n == 0) return 0;
    hipLaunchKernelGGL(cell_keys_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *desc, *sp, *in,
                       reinterpret_cast<uint32_t*>(keys), n);
    PIC_CHECK_LAUNCH();
    return 0;
}

int64_t pic_sort_workspace_bytes(int64_t n) {
    if (n <= 0) return 0;
    return (int64_t)sort_storage(n, 32) + 256;
}

int pic_sort_permutation(const int32_t* keys, int32_t* keys_sorted, int32_t* perm, int64_t n, int32_t key_bits,
                         void* workspace, int64_t workspace_bytes, void* stream) {
    PIC_REQUIRE(n >= 0 && key_bits >= 1 && key_bits <= 32);
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    if (n == 0) return 0;
    PIC_REQUIRE(keys && keys_sorted && perm);
    size_t need = sort_storage(n, key_bits);
    PIC_REQUIRE(workspace != nullptr && (size_t)workspace_bytes >= need);
    size_t bytes = need;
    hipError_t e = rocprim::radix_sort_pairs(workspace, bytes, reinterpret_cast<const uint32_t*>(keys),
                                             reinterpret_cast<uint32_t*>(keys_sorted),
                                             rocprim::counting_iterator<int32_t>(0), perm, (size_t)n, 0u,
                                             (unsigned)key_bits, as_stream(stream));
    if (e != hipSuccess) return (int)e;
    PIC_CHECK_LAUNCH();
    return 0;
}

int pic_permute_particles(const PicParticles* in, PicParticles* out, const int32_t* perm, int64_t n, void* stream) {
    PIC_REQUIRE(in && out && n >= 0);
    if (n == 0) return 0;
    PIC_REQUIRE(perm);
    for (int a = 0; a < 3; ++a) PIC_REQUIRE(in->x[a] && in->u[a] && out->x[a] && out->u[a] && in->x[a] != out->x[a]);
    PIC_REQUIRE(in->active && out->active && in->active != out->active);
    PIC_REQUIRE(n < ((int64_t)1 << 31));
    hipLaunchKernelGGL(permute_kernel, dim3(grid_1d(n, kBlock)), dim3(kBlock), 0, as_stream(stream), *in, *out, perm,
                       n);
    PIC_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
