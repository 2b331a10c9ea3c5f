// Particle cell lists on gfx950.
//
// The reference asks the application for a per-cell particle iterator
// (getIterator(cell), example/example.cpp:148-228) and walks it in
// ParticleRenderKernel (lib/isaac/isaac_particle_kernel.hpp:32-114).  This
// library builds those lists itself from raw particles.  Every frame builds only
// the render records of the visible particles (vis_bin, scan, vis_scatter,
// vis_sort below); the full cell list of all particles, on request:
//
//   1. bin_count   per particle: cell = clamp(floor(p / cw)), rel = p / cw - cell,
//                  TF index of the functor-applied attribute; ONE returning 32-bit
//                  atomicAdd per particle counts it, and its return value is the
//                  particle's arrival rank inside its cell; a particle whose TF alpha
//                  is >= 0.5 (visible) also flags its cell with a plain store of 1
//   2. scan        exclusive prefix sum of the counts: the cell offsets
//   3. bin_scatter perm[offset[cell] + rank] = i, plain stores
//   4. bin_sort    per cell: sort the cell's particle indices ascending, which makes
//                  the permutation equal to a stable counting sort by cell
//                  (bit-exact), then, in flagged cells, write the visible render
//                  records in that order from the cell's offset on; the cell's
//                  record range [offset, offset + visible) goes to vrange
//
// A particle whose TF alpha is below 0.5 can never produce a hit in the render
// kernel (the alpha test follows the geometric test and changes nothing when
// it fails), so dropping it from the render records leaves every pixel
// unchanged while removing its sphere tests.
#include "kernels.hpp"

namespace isaac_hip
{
    __device__ inline float apply_chain_p(const ChainDesc& c, const float* e, int n)
    {
        float v[4] = {e[0], n > 1 ? e[1] : 0.f, n > 2 ? e[2] : 0.f, n > 3 ? e[3] : 0.f};
#pragma unroll
        for(int k = 0; k < kMaxFunctors; k++)
        {
            const int op = c.op[k];
            if(op == OP_ADD)
            {
                for(int i = 0; i < n; i++)
                    v[i] = v[i] + c.p[k][i];
            }
            else if(op == OP_MUL)
            {
                for(int i = 0; i < n; i++)
                    v[i] = v[i] * c.p[k][i];
            }
            else if(op == OP_LENGTH)
            {
                float d;
                if(n == 1)
                    d = v[0] * v[0];
                else if(n == 2)
                    d = v[0] * v[0] + v[1] * v[1];
                else if(n == 3)
                    d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                else
                    d = (v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3]);
                v[0] = sqrtf(d);
                n = 1;
            }
            else if(op == OP_POW)
            {
                for(int i = 0; i < n; i++)
                    v[i] = powf(v[i], c.p[k][i]);
            }
            else if(op == OP_SUM)
            {
                float s = 0.0f;
                for(int i = 0; i < n; i++)
                    s = s + v[i];
                v[0] = s;
                n = 1;
            }
        }
        return v[0];
    }

    __device__ inline int cell_coord(float p, float cw, int n, float& q)
    {
        q = p / cw;
        int c = int(floorf(q));
        return iclamp(c, 0, n - 1);
    }

    __global__ __launch_bounds__(256) void bin_count_kernel(const BinParams P, const BinBuffers B)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += int64_t(gridDim.x) * blockDim.x)
        {
            float qx, qy, qz;
            const int cx = cell_coord(P.pos[3 * i + 0], P.cw[0], P.grid[0], qx);
            const int cy = cell_coord(P.pos[3 * i + 1], P.cw[1], P.grid[1], qy);
            const int cz = cell_coord(P.pos[3 * i + 2], P.cw[2], P.grid[2], qz);
            const int c = cx + P.grid[0] * (cy + P.grid[1] * cz);
            const float v = apply_chain_p(P.chain, P.attr + i * P.dim, P.dim);
            const int li = tf_index_of(v * float(P.T), P.T);
            const bool visible = !P.filter || P.lut[li].w >= 0.5f;
            B.cellOf[i] = c;
            B.rel[i] = make_float4(qx - float(cx), qy - float(cy), qz - float(cz), __int_as_float(visible ? li : -1));
            // all particles: a returning 32-bit add (the arrival rank); visible ones flag their
            // cell in the upper half of the counter buffer with a plain store of 1 (every
            // writer stores the same value, so no read-modify-write is needed)
            uint32_t* cnt = reinterpret_cast<uint32_t*>(B.count64);
            B.rank[i] = atomicAdd(cnt + c, 1u);
            if(visible)
                __hip_atomic_store(cnt + B.ncells + c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // ---- min/max request over particles (MinMaxParticleKernel,
    // lib/isaac/isaac_min_max_kernel.hpp:60-97): the reference visits cell columns
    // (x + y*gx) in the host loop's order, z inside a column, and a cell's particles
    // in list order, which for these cell lists is ascending particle index: the key
    // is ((x + y*gx) * gz + z) * 2^32 + index + 1.
    __global__ __launch_bounds__(256) void min_max_particle_kernel(const MinMaxParticleParams P, MinMaxPartial* partials)
    {
        MinMaxPartial a{FLT_MAX, -FLT_MAX, 0, 0};
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += int64_t(gridDim.x) * blockDim.x)
        {
            float q;
            const int cx = cell_coord(P.pos[3 * i + 0], P.cw[0], P.grid[0], q);
            const int cy = cell_coord(P.pos[3 * i + 1], P.cw[1], P.grid[1], q);
            const int cz = cell_coord(P.pos[3 * i + 2], P.cw[2], P.grid[2], q);
            const float v = apply_chain_p(P.chain, P.attr + i * P.dim, P.dim);
            const uint64_t cell = (uint64_t(cx) + uint64_t(cy) * uint64_t(P.grid[0])) * uint64_t(P.grid[2]) + uint64_t(cz);
            const uint64_t key = (cell << 32) + uint64_t(i) + 1;
            mm_take_min(a.vmin, a.kmin, v, key);
            mm_take_max(a.vmax, a.kmax, v, key);
        }
        a = mm_block_reduce(a);
        if(threadIdx.x == 0)
            partials[blockIdx.x] = a;
    }

    hipError_t launch_min_max_particles(const MinMaxParticleParams& P, MinMaxPartial* partials, float2* out, hipStream_t s)
    {
        if(P.n < 0 || P.n >= (int64_t(1) << 32) - 1)
            return hipErrorInvalidValue;
        if(P.n == 0)
            return launch_min_max_final(partials, 0, out, s);
        const int groups = int(std::min<int64_t>((P.n + 255) / 256, kMinMaxGroups));
        hipLaunchKernelGGL(min_max_particle_kernel, dim3(groups), dim3(256), 0, s, P, partials);
        hipError_t e = hipGetLastError();
        return e != hipSuccess ? e : launch_min_max_final(partials, groups, out, s);
    }

    // ---- exclusive scan of the per-cell counts (64-bit sums, 3-phase, 1024-wide blocks)
    constexpr int kScanBlock = 1024;

    __device__ inline unsigned long long block_inclusive_scan(unsigned long long v, unsigned long long* sh)
    {
        const int tid = threadIdx.x;
        sh[tid] = v;
        __syncthreads();
        for(int off = 1; off < kScanBlock; off <<= 1)
        {
            unsigned long long t = (tid >= off) ? sh[tid - off] : 0ull;
            __syncthreads();
            sh[tid] += t;
            __syncthreads();
        }
        return sh[tid];
    }

    // cell i's particle count (the first half of the counter buffer)
    __device__ inline unsigned long long cell_counts(const unsigned long long* in, int64_t n, int64_t i)
    {
        const uint32_t* c = reinterpret_cast<const uint32_t*>(in);
        return i < n ? static_cast<unsigned long long>(c[i]) : 0ull;
    }

    __global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const unsigned long long* in, int64_t n, unsigned long long* blockSums)
    {
        __shared__ unsigned long long sh[kScanBlock];
        const int64_t i = int64_t(blockIdx.x) * kScanBlock + threadIdx.x;
        const unsigned long long v = cell_counts(in, n, i);
        const unsigned long long s = block_inclusive_scan(v, sh);
        if(threadIdx.x == kScanBlock - 1)
            blockSums[blockIdx.x] = s;
    }

    // single workgroup: exclusive scan of the block sums, in place
    __global__ __launch_bounds__(kScanBlock) void scan_blocksums_kernel(unsigned long long* blockSums, int nb)
    {
        __shared__ unsigned long long sh[kScanBlock];
        unsigned long long carry = 0;
        for(int base = 0; base < nb; base += kScanBlock)
        {
            const int i = base + threadIdx.x;
            const unsigned long long v = i < nb ? blockSums[i] : 0ull;
            const unsigned long long s = block_inclusive_scan(v, sh);
            if(i < nb)
                blockSums[i] = carry + s - v;
            __syncthreads();
            carry += sh[kScanBlock - 1];
            __syncthreads();
        }
    }

    // writes per-cell counts (ncells) and cell offsets (ncells + 1)
    __global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(
        const unsigned long long* in,
        int64_t n,
        const unsigned long long* blockSums,
        int32_t* count,
        int32_t* offset)
    {
        __shared__ unsigned long long sh[kScanBlock];
        const int64_t i = int64_t(blockIdx.x) * kScanBlock + threadIdx.x;
        const unsigned long long v = cell_counts(in, n, i);
        const unsigned long long s = block_inclusive_scan(v, sh);
        const unsigned long long base = blockSums[blockIdx.x];
        if(i < n)
        {
            const unsigned long long ex = base + s - v;
            count[i] = int32_t(uint32_t(v));
            offset[i] = int32_t(uint32_t(ex));
        }
        if(i == n - 1)
        {
            const unsigned long long tot = base + s;
            offset[n] = int32_t(uint32_t(tot));
        }
    }

    __global__ __launch_bounds__(256) void bin_scatter_kernel(const BinParams P, const BinBuffers B)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += int64_t(gridDim.x) * blockDim.x)
            B.perm[B.offset[B.cellOf[i]] + int32_t(B.rank[i])] = int32_t(i);
    }

    // one thread per cell: ascending sort of the cell's particle indices, then
    // the visible render records in that order and the cell's record range
    __global__ __launch_bounds__(256) void bin_sort_kernel(const BinParams P, const BinBuffers B, int64_t ncells)
    {
        for(int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < ncells; c += int64_t(gridDim.x) * blockDim.x)
        {
            const int k0 = B.offset[c], k1 = B.offset[c + 1];
            const int len = k1 - k0;
            int32_t* a = B.perm + k0;
            if(len <= 64)
            {
                for(int i = 1; i < len; i++)
                {
                    const int32_t key = a[i];
                    int j = i - 1;
                    while(j >= 0 && a[j] > key)
                    {
                        a[j + 1] = a[j];
                        j--;
                    }
                    a[j + 1] = key;
                }
            }
            else
            {
                // Shell sort (Ciura gaps) for crowded cells
                const int gaps[8] = {701, 301, 132, 57, 23, 10, 4, 1};
                for(int g = 0; g < 8; g++)
                {
                    const int gap = gaps[g];
                    for(int i = gap; i < len; i++)
                    {
                        const int32_t key = a[i];
                        int j = i;
                        while(j >= gap && a[j - gap] > key)
                        {
                            a[j] = a[j - gap];
                            j -= gap;
                        }
                        a[j] = key;
                    }
                }
            }
            int w = k0;
            // cells without a visible particle (flag 0) gather no records
            if(reinterpret_cast<const uint32_t*>(B.count64)[B.ncells + c] != 0u)
            {
                for(int k = 0; k < len; k++)
                {
                    const int32_t i = a[k];
                    const float4 r = B.rel[i];
                    if(__float_as_int(r.w) >= 0)
                    {
                        B.recs[w] = r;
                        if(B.recRadius)
                            B.recRadius[w] = P.radius[i];
                        w++;
                    }
                }
            }
            B.vrange[c] = make_int2(k0, w);
        }
    }

    // ---- the visible-records build: the render needs, per cell, only the records of
    // the particles whose transfer-function alpha is >= 0.5, in ascending particle
    // index (the reference's iteration order decides ties).  Invisible particles cost
    // one attribute read and one bit.
    //
    // step 1, per particle: the TF index of its functor-applied attribute; each wave
    // takes 4 x 64 consecutive particles (all attribute reads, then all transfer
    // lookups in flight together) and stores 4 words of 64 visibility bits; a visible
    // particle counts itself in its cell (a returning add: its arrival rank) and keeps
    // cell, rank and record at its index
    constexpr int kVisPerLane = 4;
    __global__ __launch_bounds__(256) void vis_bin_kernel(const BinParams P, const BinBuffers B)
    {
        uint32_t* cnt = reinterpret_cast<uint32_t*>(B.count64);
        const uint32_t lane = threadIdx.x & 63u;
        const int64_t waves = int64_t(gridDim.x) * (blockDim.x >> 6);
        for(int64_t w = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); w * 64 * kVisPerLane < P.n; w += waves)
        {
            const int64_t base = w * 64 * kVisPerLane;
            float v[kVisPerLane];
#pragma unroll
            for(int k = 0; k < kVisPerLane; k++)
            {
                const int64_t i = base + k * 64 + lane;
                v[k] = i < P.n ? apply_chain_p(P.chain, P.attr + i * P.dim, P.dim) : 0.0f;
            }
            int li[kVisPerLane];
            bool visible[kVisPerLane];
#pragma unroll
            for(int k = 0; k < kVisPerLane; k++)
            {
                const int64_t i = base + k * 64 + lane;
                li[k] = tf_index_of(v[k] * float(P.T), P.T);
                visible[k] = i < P.n && (!P.filter || P.lut[li[k]].w >= 0.5f);
            }
#pragma unroll
            for(int k = 0; k < kVisPerLane; k++)
            {
                const int64_t i0 = base + k * 64;
                const unsigned long long mask = __ballot(visible[k]);
                if(lane == 0u && i0 < P.n)
                    B.visMask[i0 >> 6] = mask;
                if(visible[k])
                {
                    const int64_t i = i0 + lane;
                    float qx, qy, qz;
                    const int cx = cell_coord(P.pos[3 * i + 0], P.cw[0], P.grid[0], qx);
                    const int cy = cell_coord(P.pos[3 * i + 1], P.cw[1], P.grid[1], qy);
                    const int cz = cell_coord(P.pos[3 * i + 2], P.cw[2], P.grid[2], qz);
                    const int c = cx + P.grid[0] * (cy + P.grid[1] * cz);
                    B.cellOf[i] = c;
                    B.rank[i] = atomicAdd(cnt + c, 1u);
                    B.rel[i] = make_float4(qx - float(cx), qy - float(cy), qz - float(cz), __int_as_float(li[k]));
                }
            }
        }
    }

    // step 3 (after the scan of the visible counts): each visible particle's index at
    // its cell's offset plus its rank
    __global__ __launch_bounds__(256) void vis_scatter_kernel(const BinParams P, const BinBuffers B)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P.n; i += int64_t(gridDim.x) * blockDim.x)
            if((B.visMask[i >> 6] >> (i & 63)) & 1ull)
                B.perm[B.offset[B.cellOf[i]] + int32_t(B.rank[i])] = int32_t(i);
    }

    // step 4, per cell: its visible particles in ascending index, then their records
    // and the cell's record range
    __global__ __launch_bounds__(256) void vis_sort_kernel(const BinParams P, const BinBuffers B, int64_t ncells)
    {
        for(int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < ncells; c += int64_t(gridDim.x) * blockDim.x)
        {
            const int k0 = B.offset[c], k1 = B.offset[c + 1];
            int32_t* a = B.perm + k0;
            const int len = k1 - k0;
            // insertion sort, with Shell sort's Ciura gaps first for crowded cells
            const int gaps[8] = {701, 301, 132, 57, 23, 10, 4, 1};
            for(int g = len > 64 ? 0 : 7; g < 8; g++)
            {
                const int gap = gaps[g];
                for(int i = gap; i < len; i++)
                {
                    const int32_t key = a[i];
                    int j = i;
                    while(j >= gap && a[j - gap] > key)
                    {
                        a[j] = a[j - gap];
                        j -= gap;
                    }
                    a[j] = key;
                }
            }
            for(int k = 0; k < len; k++)
            {
                const int32_t i = a[k];
                B.recs[k0 + k] = B.rel[i];
                if(B.recRadius)
                    B.recRadius[k0 + k] = P.radius[i];
            }
            B.vrange[c] = make_int2(k0, k1);
        }
    }

    __global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* p, uint32_t value, size_t n)
    {
        for(size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
            p[i] = value;
    }

    hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s)
    {
        if(n == 0)
            return hipSuccess;
        size_t blocks = (n + 255) / 256;
        if(blocks > 8192)
            blocks = 8192;
        hipLaunchKernelGGL(fill_u32_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, p, value, n);
        return hipGetLastError();
    }

    size_t bin_scan_workspace_elems(int64_t ncells)
    {
        return size_t((ncells + kScanBlock - 1) / kScanBlock) + 64;
    }

    hipError_t launch_bin(const BinParams& P, const BinBuffers& B, hipStream_t s)
    {
        const int64_t ncells = int64_t(P.grid[0]) * P.grid[1] * P.grid[2];
        hipError_t e;
        if((e = launch_fill_u32(reinterpret_cast<uint32_t*>(B.count64), 0u, size_t(ncells) * 2, s)) != hipSuccess)
            return e;
        int64_t pb = (P.n + 255) / 256;
        if(pb > 8192)
            pb = 8192;
        if(P.n > 0)
            hipLaunchKernelGGL(bin_count_kernel, dim3(unsigned(pb)), dim3(256), 0, s, P, B);
        const int nb = int((ncells + kScanBlock - 1) / kScanBlock);
        hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanBlock), 0, s, B.count64, ncells, B.scanTmp);
        hipLaunchKernelGGL(scan_blocksums_kernel, dim3(1), dim3(kScanBlock), 0, s, B.scanTmp, nb);
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanBlock), 0, s, B.count64, ncells, B.scanTmp, B.count,
                           B.offset);
        if(P.n > 0)
            hipLaunchKernelGGL(bin_scatter_kernel, dim3(unsigned(pb)), dim3(256), 0, s, P, B);
        int64_t cb = (ncells + 255) / 256;
        if(cb > 8192)
            cb = 8192;
        hipLaunchKernelGGL(bin_sort_kernel, dim3(unsigned(cb)), dim3(256), 0, s, P, B, ncells);
        return hipGetLastError();
    }

    hipError_t launch_bin_visible(const BinParams& P, const BinBuffers& B, hipStream_t s)
    {
        const int64_t ncells = int64_t(P.grid[0]) * P.grid[1] * P.grid[2];
        hipError_t e;
        if((e = launch_fill_u32(reinterpret_cast<uint32_t*>(B.count64), 0u, size_t(ncells), s)) != hipSuccess)
            return e;
        int64_t pb = (P.n + 255) / 256;
        if(pb > 8192)
            pb = 8192;
        // vis_bin: 4 waves per workgroup, 4 x 64 particles per wave
        int64_t vb = (P.n + 256 * kVisPerLane - 1) / (256 * kVisPerLane);
        if(vb > 8192)
            vb = 8192;
        if(P.n > 0)
            hipLaunchKernelGGL(vis_bin_kernel, dim3(unsigned(vb)), dim3(256), 0, s, P, B);
        const int nb = int((ncells + kScanBlock - 1) / kScanBlock);
        hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanBlock), 0, s, B.count64, ncells, B.scanTmp);
        hipLaunchKernelGGL(scan_blocksums_kernel, dim3(1), dim3(kScanBlock), 0, s, B.scanTmp, nb);
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanBlock), 0, s, B.count64, ncells, B.scanTmp, B.count,
                           B.offset);
        if(P.n > 0)
            hipLaunchKernelGGL(vis_scatter_kernel, dim3(unsigned(pb)), dim3(256), 0, s, P, B);
        int64_t cb = (ncells + 255) / 256;
        if(cb > 8192)
            cb = 8192;
        hipLaunchKernelGGL(vis_sort_kernel, dim3(unsigned(cb)), dim3(256), 0, s, P, B, ncells);
        return hipGetLastError();
    }
} // namespace isaac_hip
