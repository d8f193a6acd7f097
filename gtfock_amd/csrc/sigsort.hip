// sigsort.hip -- the significant shell pairs of one pair type, compacted
// and ordered on the device (pfock/screening.c:111-175 keeps a pair when its
// Schwarz value exceeds tol^2 / vmax of its molecule).
//
// Compaction: a wavefront ballot and prefix popcount per 64 pairs, one
// count per workgroup, an exclusive scan of the counts, then every kept pair
// is written at its rank -- the kept ids stay in id order.
//
// Order (stable radix sorts, rocPRIM):
//   1. value descending (ties by id)            -> molecule value order
//   2. molecule                                  -> rank = position within the
//                                                   molecule's value order
//   3. (molecule, primitive-count bucket)        -> segments, value order
//      inside ("sorted" lists: value, rank, primitive count)
//   4. (64-block of the segment, primitive count descending)
//                                                -> the ket-list order
// The caller copies what the host-side tiling needs.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "gtf_internal.h"
#include "screen.h"

__device__ __forceinline__ int prim_bucket(int np)
{
    return np <= 1 ? 0 : np <= 4 ? 1 : np <= 12 ? 2 : np <= 32 ? 3 : 4;
}

__device__ __forceinline__ double bits_to_double(unsigned long long b)
{
    return __longlong_as_double((long long)b);
}

// per workgroup: number of kept pairs
__global__ __launch_bounds__(256) void sig_count_kernel(int npairs, const int* __restrict__ type, int t,
                                                        const unsigned long long* __restrict__ vbits,
                                                        const int* __restrict__ pmol, const double* __restrict__ eta,
                                                        int* __restrict__ block_count)
{
    __shared__ int wcount[4];
    const int id = blockIdx.x * 256 + threadIdx.x;
    const bool keep = id < npairs && type[id] == t && bits_to_double(vbits[id]) > eta[pmol[id]];
    const unsigned long long m = __ballot(keep);
    if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) block_count[blockIdx.x] = wcount[0] + wcount[1] + wcount[2] + wcount[3];
}

// exclusive scan of the workgroup counts (one workgroup, any count)
__global__ __launch_bounds__(1024) void sig_scan_kernel(int nblocks, const int* __restrict__ block_count,
                                                        int* __restrict__ block_off, int* __restrict__ total)
{
    __shared__ int part[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblocks; b0 += 1024) {
        const int b = b0 + threadIdx.x;
        const int v = b < nblocks ? block_count[b] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int s = 1; s < 1024; s <<= 1) {
            const int add = threadIdx.x >= s ? part[threadIdx.x - s] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (b < nblocks) block_off[b] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void sig_scatter_kernel(int npairs, const int* __restrict__ type, int t,
                                                          const unsigned long long* __restrict__ vbits,
                                                          const int* __restrict__ pmol,
                                                          const double* __restrict__ eta,
                                                          const int* __restrict__ block_off, int* __restrict__ ids,
                                                          unsigned long long* __restrict__ key)
{
    __shared__ int wbase[4];
    const int id = blockIdx.x * 256 + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool keep = id < npairs && type[id] == t && bits_to_double(vbits[id]) > eta[pmol[id]];
    const unsigned long long m = __ballot(keep);
    if (lane == 0) wbase[w] = __popcll(m);
    __syncthreads();
    if (keep) {
        int base = block_off[blockIdx.x];
        for (int k = 0; k < w; k++) base += wbase[k];
        const int at = base + __popcll(m & ((1ull << lane) - 1ull));
        ids[at] = id;
        key[at] = ~vbits[id];  // Schwarz values are >= 0: descending value order
    }
}

__global__ void sig_mol_key_kernel(int n, const int* __restrict__ ids, const int* __restrict__ pmol,
                                   unsigned* __restrict__ key)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) key[k] = (unsigned)pmol[ids[k]];
}

// rank within the molecule's value order; segment key (molecule, bucket)
__global__ void sig_rank_kernel(int n, const int* __restrict__ ids, const unsigned* __restrict__ molkey,
                                const int* __restrict__ nprim, int nbucket, int* __restrict__ rank,
                                unsigned* __restrict__ segkey, int* __restrict__ pos)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    // first position of this molecule: binary search over the sorted keys
    const unsigned m = molkey[k];
    int lo = 0, hi = k;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (molkey[mid] < m) lo = mid + 1;
        else hi = mid;
    }
    rank[k] = k - lo;
    segkey[k] = m * nbucket + prim_bucket(nprim[ids[k]]);
    pos[k] = k;
}

// the segment lists in value order, and each segment's start and length
__global__ void sig_segment_kernel(int n, const unsigned* __restrict__ segkey, const int* __restrict__ pos,
                                   const int* __restrict__ ids, const int* __restrict__ rank,
                                   const unsigned long long* __restrict__ vbits, const int* __restrict__ nprim,
                                   int* __restrict__ s_id, double* __restrict__ s_val, int* __restrict__ s_rank,
                                   int* __restrict__ s_nprim, int* __restrict__ seg_start, int* __restrict__ seg_len)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int id = ids[pos[k]];
    s_id[k] = id;
    s_val[k] = bits_to_double(vbits[id]);
    s_rank[k] = rank[pos[k]];
    s_nprim[k] = nprim[id];
    const unsigned sg = segkey[k];
    if (k == 0 || segkey[k - 1] != sg) seg_start[sg] = k;
    if (k == n - 1 || segkey[k + 1] != sg) {
        int lo = 0, hi = k;  // the segment's first position
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (segkey[mid] < sg) lo = mid + 1;
            else hi = mid;
        }
        seg_len[sg] = k + 1 - lo;
    }
}

// key of the ket-list order: the 64-block of the segment, then primitive
// count descending
__global__ void sig_block_key_kernel(int n, const unsigned* __restrict__ segkey, const int* __restrict__ seg_start,
                                     const int* __restrict__ s_nprim, unsigned long long* __restrict__ key,
                                     int* __restrict__ pos)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int s0 = seg_start[segkey[k]];
    const long blk = s0 + ((long)(k - s0) / GTF_KET_BLOCK) * GTF_KET_BLOCK;
    key[k] = (unsigned long long)blk * 256ull + (unsigned long long)(255 - min(255, s_nprim[k]));
    pos[k] = k;
}

__global__ void sig_final_kernel(int n, const int* __restrict__ pos, const int* __restrict__ s_id,
                                 const double* __restrict__ s_val, const int* __restrict__ s_rank,
                                 const int* __restrict__ pmol, int* __restrict__ sig, double* __restrict__ sigv,
                                 int* __restrict__ rank, int* __restrict__ sigmol)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int p = pos[k];
    sig[k] = s_id[p];
    sigv[k] = s_val[p];
    rank[k] = s_rank[p];
    sigmol[k] = pmol[s_id[p]];
}

static inline int nblk(long n) { return (int)((n + 255) / 256); }

int gtf_sig_count(hipStream_t st, int npairs, const int* type, int t, const unsigned long long* vbits, const int* pmol,
                  const double* eta, int* block_count, int* block_off, int* total)
{
    const int nb = nblk(npairs);
    sig_count_kernel<<<nb, 256, 0, st>>>(npairs, type, t, vbits, pmol, eta, block_count);
    GTF_CHECK(hipGetLastError());
    sig_scan_kernel<<<1, 1024, 0, st>>>(nb, block_count, block_off, total);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_sig_order(hipStream_t st, int npairs, const int* type, int t, const unsigned long long* vbits,
                  const int* pmol, const double* eta, const int* nprim, int nbucket, const int* block_off, int n,
                  SigWork& w, SigOut& o)
{
    if (n <= 0) return 0;
    const int nb = nblk(npairs), nk = nblk(n);
    sig_scatter_kernel<<<nb, 256, 0, st>>>(npairs, type, t, vbits, pmol, eta, block_off, w.ids_a, w.k64_a);
    GTF_CHECK(hipGetLastError());
    size_t bytes = w.temp_bytes;
    // 1. value descending, ties by id (stable over the id order)
    GTF_CHECK(rocprim::radix_sort_pairs(w.temp, bytes, w.k64_a, w.k64_b, w.ids_a, w.ids_b, n, 0, 64, st));
    // 2. molecule
    sig_mol_key_kernel<<<nk, 256, 0, st>>>(n, w.ids_b, pmol, w.k32_a);
    GTF_CHECK(hipGetLastError());
    bytes = w.temp_bytes;
    GTF_CHECK(rocprim::radix_sort_pairs(w.temp, bytes, w.k32_a, w.k32_b, w.ids_b, w.ids_a, n, 0, 32, st));
    // ranks, and the (molecule, bucket) key
    sig_rank_kernel<<<nk, 256, 0, st>>>(n, w.ids_a, w.k32_b, nprim, nbucket, w.rank, w.k32_a, w.pos_a);
    GTF_CHECK(hipGetLastError());
    // 3. segments
    bytes = w.temp_bytes;
    GTF_CHECK(rocprim::radix_sort_pairs(w.temp, bytes, w.k32_a, w.k32_b, w.pos_a, w.pos_b, n, 0, 32, st));
    sig_segment_kernel<<<nk, 256, 0, st>>>(n, w.k32_b, w.pos_b, w.ids_a, w.rank, vbits, nprim, o.s_id, o.sorted,
                                           o.sortrank, o.sortnprim, o.seg_start, o.seg_len);
    GTF_CHECK(hipGetLastError());
    // 4. 64-blocks by primitive count
    sig_block_key_kernel<<<nk, 256, 0, st>>>(n, w.k32_b, o.seg_start, o.sortnprim, w.k64_a, w.pos_a);
    GTF_CHECK(hipGetLastError());
    bytes = w.temp_bytes;
    GTF_CHECK(rocprim::radix_sort_pairs(w.temp, bytes, w.k64_a, w.k64_b, w.pos_a, w.pos_b, n, 0, 64, st));
    sig_final_kernel<<<nk, 256, 0, st>>>(n, w.pos_b, o.s_id, o.sorted, o.sortrank, pmol, o.sig, o.sigv, o.rank,
                                         o.sigmol);
    GTF_CHECK(hipGetLastError());
    return 0;
}

size_t gtf_sig_temp_bytes(int n)
{
    size_t b64 = 0, b32 = 0;
    unsigned long long* k64 = nullptr;
    unsigned* k32 = nullptr;
    int* v = nullptr;
    (void)rocprim::radix_sort_pairs(nullptr, b64, k64, k64, v, v, n, 0, 64);
    (void)rocprim::radix_sort_pairs(nullptr, b32, k32, k32, v, v, n, 0, 32);
    return std::max(b64, b32);
}
