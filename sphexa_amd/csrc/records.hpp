// Per-particle neighbour records for the pair loops.
//
// A pair loop reads, for every neighbour j of its lane's target, a handful of fields of j.
// Read from the SoA Dataset arrays, every field is a separate 64-lane gather touching tens of
// cache lines; the loops were bound by the vector-memory address path, not by ALU or HBM.
// Before each pair loop a coalesced pack kernel therefore writes exactly the fields that loop
// needs for j into one 16-byte aligned record per particle (workspace memory, SFC order), so a
// neighbour costs a few 16-byte loads from one or two cache lines. Positions are FP64 in the plain
// form and 32-bit fixed point in the FP32 form (Coord, pairs.hpp). Derived per-particle
// quantities (1/h, m/rho, xm/kx, kx m/xm) are formed in the pack with the same expression the
// pair loop would evaluate, so results are unchanged.
#pragma once

#include <cstring>

#include "pairs.hpp"

namespace sphx
{

//! XMass / std density: position and mass
template<class T>
struct alignas(16) RecMass
{
    Coord<T> x, y, z;
    T      m;
};

//! VeDefGradh: position, mass, volume-element mass xm
template<class T>
struct alignas(16) RecGradh
{
    Coord<T> x, y, z;
    T      m, xm;
};

//! std IAD: position and volume m/rho
template<class T>
struct alignas(16) RecVol
{
    Coord<T> x, y, z;
    T      vol;
};

//! VE IAD + div/curl v: position, volume xm/kx, xm, velocity
template<class T>
struct alignas(16) RecIadVe
{
    Coord<T> x, y, z;
    T      vol, xm, vx, vy, vz;
};

//! AV switches: position, volume xm/kx, sound speed, div v, velocity
template<class T>
struct alignas(16) RecAv
{
    Coord<T> x, y, z;
    T      vol, c, divv, vx, vy, vz;
};

//! std momentum-energy: vol = m/rho, pro = p/rho
template<class T>
struct alignas(16) RecMomStd
{
    Coord<T> x, y, z;
    T      vx, vy, vz, hInv, m, vol, pro, c, c11, c12, c13, c22, c23, c33;
};

//! VE momentum-energy: rho = kx m/xm, vol = m/rho (the AV-cleaning gradients stay in SoA)
template<class T>
struct alignas(16) RecMomVe
{
    Coord<T> x, y, z;
    T      vx, vy, vz, hInv, m, c, rho, vol, xm, alpha, prho, c11, c12, c13, c22, c23, c33;
};

// 16-byte loads per neighbour, FP32 form: 1 (mass, volume), 2 (grad-h, VE IAD), 3 (AV), 5 (momentum)
static_assert(sizeof(RecMass<float>) == 16 && sizeof(RecVol<float>) == 16, "record layout");
static_assert(sizeof(RecGradh<float>) == 32 && sizeof(RecIadVe<float>) == 32 && sizeof(RecAv<float>) == 48,
              "record layout");
static_assert(sizeof(RecMomStd<float>) == 80 && sizeof(RecMomVe<float>) == 80, "record layout");

//! the VeDefGradh record, from the pack kernel or the XMass epilogue (the same bits)
template<class T>
__device__ __forceinline__ RecGradh<T> gradhRecord(Coord<T> x, Coord<T> y, Coord<T> z, T m, T xm)
{
    RecGradh<T> r;
    r.x = x, r.y = y, r.z = z;
    r.m = m, r.xm = xm;
    return r;
}

//! the std IAD record, from the pack kernel or the density epilogue (the same bits)
template<class T>
__device__ __forceinline__ RecVol<T> volRecord(Coord<T> x, Coord<T> y, Coord<T> z, T m, T rho)
{
    RecVol<T> r;
    r.x = x, r.y = y, r.z = z;
    r.vol = fastDiv(m, rho);
    return r;
}

//! the VE IAD record, from the pack kernel or the VeDefGradh epilogue (the same bits)
template<class T>
__device__ __forceinline__ RecIadVe<T> iadVeRecord(Coord<T> x, Coord<T> y, Coord<T> z, T xm, T kx, T vx, T vy, T vz)
{
    RecIadVe<T> r;
    r.x = x, r.y = y, r.z = z;
    r.xm  = xm;
    r.vol = fastDiv(xm, kx);
    r.vx = vx, r.vy = vy, r.vz = vz;
    return r;
}

//! workspace buffer `buf` (0 or 1) for `n` records of type R
template<class R>
R* recordBuffer(sphx_workspace* ws, size_t n, int buf = 0)
{
    auto& b = buf ? ws->records2 : ws->records;
    b.reserve((n * sizeof(R) + 15) / 16);
    return reinterpret_cast<R*>(b.get());
}

//! record kinds a pair loop's epilogue can leave for the next call (SPHX_FLAG_CHAIN)
enum ChainKind
{
    kChainAv    = 1, // sphx_iad_divv_curlv -> sphx_av_switches
    kChainMomVe = 2, // sphx_av_switches -> sphx_momentum_energy_ve
    kChainGradh = 3, // sphx_xmass -> sphx_ve_def_gradh
    kChainMomStd = 4, // sphx_iad_std -> sphx_momentum_energy_std
    kChainIadVe  = 5, // sphx_ve_def_gradh -> (sphx_eos_ve) -> sphx_iad_divv_curlv
    kChainVol    = 6  // sphx_density_std -> (sphx_eos_std) -> sphx_iad_std
};

//! the call takes part in record chaining: flag set, the full particle range, consecutive target groups
inline bool chainable(const PairArgs& a, const sphx_params_t* p)
{
    return (p->flags & SPHX_FLAG_CHAIN) && a.L.groups.first == 0 && a.L.groups.last == a.numParticles &&
           a.L.groups.start == nullptr;
}

//! the record buffer holding the records of `kind` that the previous call wrote for this one, or -1
inline int takeChain(sphx_workspace* ws, int kind, const PairArgs& a, const sphx_fields_t* f, const sphx_params_t* p)
{
    auto& c = ws->chain;
    // the producing call directly before this one, or (IAD records only) with one EOS of the same
    // formulation on the same fields between: that EOS writes none of the fields the records hold
    bool next    = c.serial + 1 == ws->callSerial;
    bool eosKind = (kind == kChainIadVe && ws->eosKind == 2) || (kind == kChainVol && ws->eosKind == 1);
    bool overEos = eosKind && c.serial + 2 == ws->callSerial && ws->eosSerial == c.serial + 1 &&
                   std::memcmp(&ws->eosFields, f, sizeof(*f)) == 0;
    bool ok = chainable(a, p) && c.kind == kind && (next || overEos) && c.n == a.numParticles &&
              c.form == p->hydro_type && std::memcmp(&c.fields, f, sizeof(*f)) == 0;
    for (int k = 0; k < 3; ++k)
    {
        ok = ok && c.lo[k] == a.box.lo[k] && c.hi[k] == a.box.hi[k] && c.pbc[k] == a.box.pbc[k];
    }
    c.kind = 0;
    return ok ? c.buf : -1;
}

inline void leaveChain(sphx_workspace* ws, int kind, int buf, const PairArgs& a, const sphx_fields_t* f,
                const sphx_params_t* p)
{
    auto& c  = ws->chain;
    c.kind   = kind;
    c.serial = ws->callSerial;
    c.n      = a.numParticles;
    c.form   = p->hydro_type;
    c.buf    = buf;
    c.fields = *f;
    for (int k = 0; k < 3; ++k)
    {
        c.lo[k] = a.box.lo[k], c.hi[k] = a.box.hi[k], c.pbc[k] = a.box.pbc[k];
    }
}

} // namespace sphx
