// filt_zmarch.hpp -- pieces shared by the z-marching tile kernels
// (filt_stencil.hip, filt_step.hip): the LDS tile layout with its one-cell
// ring, the ring cell of each thread, and loads / stores through 32-bit byte
// offsets from uniform array bases (one VGPR per address).
#pragma once

#include <hip/hip_runtime.h>

namespace filt {
namespace zm {

template <int TX, int TY>
struct Lds {
    static constexpr int SX = TX + 2;
    static constexpr int SN = SX * (TY + 2);
    __device__ static inline int at(int a, int y, int x) { return a * SN + y * SX + x; }
};

__device__ __forceinline__ double ld(const double* base, unsigned off)
{
    return *(const double*)((const char*)base + off);
}

__device__ __forceinline__ void st(double* base, unsigned off, double v)
{
    *(double*)((char*)base + off) = v;
}

// Store with the non-temporal hint: outputs that no later access of the
// same launch reads stream out instead of displacing the L2 lines that
// neighbouring tiles read as their halo (measured on the fused step: 7%).
__device__ __forceinline__ void stn(double* base, unsigned off, double v)
{
    __builtin_nontemporal_store(v, (double*)((char*)base + off));
}

__device__ __forceinline__ int ldi(const int* base, unsigned off4)
{
    return *(const int*)((const char*)base + off4);
}

// Buffer view of an output array of `bytes` bytes.  A buffer store is range
// checked by the hardware: a lane whose offset lies past the array stores
// nothing.  A plane loop's stores are therefore issued by every lane, with
// an offset of DROP where the lane has nothing to store, and no branch skips
// them.  That keeps the number of memory operations issued after a load the
// same on every path through the loop, and with it the compiler's wait for
// the load exact: a store that some path skips makes the wait cover every
// memory operation outstanding, the plane's stores included.
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef __attribute__((ext_vector_type(2))) unsigned U32x2;
constexpr unsigned DROP = 0xFFFFFFF8u;  // past every array (arrays stay below 4 GiB)
constexpr int BUF_WORD3 = 0x00020000;   // raw buffer, 32-bit data format
constexpr int VMCNT0 = 0x0F70;          // s_waitcnt vmcnt(0), other counters free

__device__ __forceinline__ Rsrc out_rsrc(double* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, BUF_WORD3);
}

// Non-temporal buffer store of one double.
__device__ __forceinline__ void bstn(Rsrc r, unsigned off, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U32x2, v), r, (int)off, 0,
                                          2 /* non-temporal */);
}

// The halo ring of a TX x TY tile without its corners (a cross stencil never
// reads them): 2*TX + 2*TY cells, cell h at LDS row ly, column lx and at
// (di, dj) from the tile origin.  Rows y-1 and y+TY first, then columns x-1
// and x+TX.
template <int TX, int TY>
struct Ring {
    static constexpr int N = 2 * TX + 2 * TY;
    __device__ static inline void cell(int h, int& ly, int& lx, int& di, int& dj)
    {
        if (h < TX) {
            ly = 0, lx = h + 1, di = h, dj = -1;
        } else if (h < 2 * TX) {
            ly = TY + 1, lx = h - TX + 1, di = h - TX, dj = TY;
        } else if (h < 2 * TX + TY) {
            ly = h - 2 * TX + 1, lx = 0, di = -1, dj = ly - 1;
        } else {
            ly = h - 2 * TX - TY + 1, lx = TX + 1, di = TX, dj = ly - 1;
        }
    }
};

// Thread (tx, ty)'s ring cell, when it has one inside the grid: its LDS
// slot (array 0) and its column's byte offset.
template <int TX, int TY>
__device__ inline bool ring_slot(int tx, int ty, int i0, int j0, int nx, int ny, unsigned sy,
                                 int& slot, unsigned& hcol)
{
    using R = Ring<TX, TY>;
    static_assert(R::N <= TX * TY, "one ring cell per thread at most");
    const int h = tx + TX * ty;
    slot = 0;
    hcol = 0;
    if (h >= R::N)
        return false;
    int ly, lx, di, dj;
    R::cell(h, ly, lx, di, dj);
    const int gi = i0 + di, gj = j0 + dj;
    if (gi < 0 || gi >= nx || gj < 0 || gj >= ny)
        return false;
    slot = Lds<TX, TY>::at(0, ly, lx);
    hcol = 8u * gi + sy * gj;
    return true;
}

}  // namespace zm
}  // namespace filt
