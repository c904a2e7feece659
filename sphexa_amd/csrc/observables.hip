// Conserved quantities of one rank's particles: kinetic and internal energy, linear and angular
// momentum, total neighbour count, in FP64, as conservedQuantitiesGpu computes them
// (main/src/observables/conserved_gpu.cu:50-95). Deterministic two-stage reduction: a grid-stride
// loop per thread, block sums through LDS into per-block partials, then one block sums the
// partials in a fixed order (no atomics), so repeated calls agree bit for bit.
#include "pairs.hpp"

namespace sphx
{

namespace
{

constexpr unsigned kRedThreads = 256;
constexpr unsigned kRedBlocks  = 1024;
constexpr int      kQuantities = 9; // eKin (m |v|^2), m u, linmom[3], angmom[3], nc

//! sum of v over the block into out[0..kQuantities) (every thread passes its partials)
__device__ __forceinline__ void blockSum(double (&v)[kQuantities], double* out)
{
    __shared__ double part[kRedThreads / kWave][kQuantities];
    for (int q = 0; q < kQuantities; ++q)
        v[q] = waveSum(v[q]);
    if (laneId() == 0)
        for (int q = 0; q < kQuantities; ++q)
            part[threadIdx.x / kWave][q] = v[q];
    __syncthreads();
    if (threadIdx.x < kQuantities)
    {
        double s = 0;
        for (unsigned w = 0; w < kRedThreads / kWave; ++w)
            s += part[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

/* EMom (conserved_gpu.cu:50-69) per particle: m |v|^2, m v, m (x cross v) with v and m promoted to
 * FP64; and m u (or m temp, scaled by cv afterwards: inner_product at :86-89) */
template<class T>
__global__ void __launch_bounds__(kRedThreads) conservedKernel(uint32_t first, uint32_t last, const double* x,
                                                              const double* y, const double* z, const T* vx,
                                                              const T* vy, const T* vz, const T* m, const double* e,
                                                              const uint32_t* nc, double* partial)
{
    double v[kQuantities] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x; i < last; i += gridDim.x * blockDim.x)
    {
        const double X[3] = {x[i], y[i], z[i]};
        const double V[3] = {double(vx[i]), double(vy[i]), double(vz[i])};
        const double mi   = double(m[i]);
        v[0] += mi * (V[0] * V[0] + (V[1] * V[1] + V[2] * V[2])); // norm2: right fold (util/array.hpp:237-240)
        if (e) { v[1] += mi * e[i]; }
        for (int d = 0; d < 3; ++d)
            v[2 + d] += mi * V[d];
        v[5] += mi * (X[1] * V[2] - X[2] * V[1]);
        v[6] += mi * (X[2] * V[0] - X[0] * V[2]);
        v[7] += mi * (X[0] * V[1] - X[1] * V[0]);
        if (nc) { v[8] += double(nc[i]); }
    }
    blockSum(v, partial + size_t(blockIdx.x) * kQuantities);
}

__global__ void __launch_bounds__(kRedThreads) finalSumKernel(const double* partial, unsigned numPartials,
                                                             double* out)
{
    double v[kQuantities] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (unsigned b = threadIdx.x; b < numPartials; b += blockDim.x)
        for (int q = 0; q < kQuantities; ++q)
            v[q] += partial[size_t(b) * kQuantities + q];
    blockSum(v, out);
}

template<class T>
void launchConserved(sphx_workspace* ws, const sphx_fields_t* f, uint32_t first, uint32_t last, const double* e,
                     double* partial, unsigned blocks)
{
    conservedKernel<T><<<blocks, kRedThreads, 0, ws->stream>>>(first, last, f->x, f->y, f->z,
                                                                static_cast<const T*>(f->vx),
                                                                static_cast<const T*>(f->vy),
                                                                static_cast<const T*>(f->vz),
                                                                static_cast<const T*>(f->m), e, f->nc, partial);
    SPHX_CHECK_LAUNCH();
}

} // namespace

void conservedQuantities(sphx_workspace* ws, const sphx_fields_t* f, const sphx_params_t* p, uint32_t first,
                         uint32_t last, double cvIn, sphx_conserved_t* out)
{
    SPHX_REQUIRE(f->x && f->y && f->z && f->vx && f->vy && f->vz && f->m, SPHX_ERR_ARG,
                 "conserved quantities need x, y, z, vx, vy, vz and m");
    SPHX_REQUIRE(first <= last, SPHX_ERR_ARG, "invalid range");
    *out = sphx_conserved_t{};
    if (first == last) { return; }
    const unsigned blocks = std::min(ceilDiv(last - first, kRedThreads), kRedBlocks);
    ws->reduce64.reserve(size_t(kRedBlocks + 1) * kQuantities);
    double*       partial = ws->reduce64.get();
    double*       total   = partial + size_t(kRedBlocks) * kQuantities;
    const double* e       = f->temp ? f->temp : f->u;
    if (p->hydro_type == SPHX_F32) { launchConserved<float>(ws, f, first, last, e, partial, blocks); }
    else { launchConserved<double>(ws, f, first, last, e, partial, blocks); }
    finalSumKernel<<<1, kRedThreads, 0, ws->stream>>>(partial, blocks, total);
    SPHX_CHECK_LAUNCH();
    double t[kQuantities];
    SPHX_HIP(hipMemcpyAsync(t, total, sizeof(t), hipMemcpyDeviceToHost, ws->stream));
    SPHX_HIP(hipStreamSynchronize(ws->stream));
    // cv times sum m temp with cv = idealGasCv(muiConst, gamma) as the reference's host forms it
    // (conserved_quantities.hpp:127-128), in the mass type
    double cv = 1.0;
    if (f->temp && cvIn >= 0) { cv = cvIn; }
    else if (f->temp)
    {
        cv = p->hydro_type == SPHX_F32 ? double(float(8.317e7f / float(p->muiConst) / (p->gamma - 1.0)))
                                       : 8.317e7 / p->muiConst / (p->gamma - 1.0);
    }
    out->eKin  = 0.5 * t[0];
    out->eInt  = cv * t[1];
    out->ncsum = t[8];
    for (int d = 0; d < 3; ++d)
    {
        out->linmom[d] = t[2 + d];
        out->angmom[d] = t[5 + d];
    }
}

} // namespace sphx
