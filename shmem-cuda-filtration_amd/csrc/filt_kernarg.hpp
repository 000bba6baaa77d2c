// filt_kernarg.hpp -- reading a kernel's struct arguments where they are used.
//
// A z-marching kernel that takes filt_fields and filt_params by value uses
// some twenty array pointers and a dozen parameters in every iteration of its
// plane loop.  The compiler loads each of them into SGPRs once and keeps them
// for the whole loop; beyond the ~100 SGPRs a wave has it spills them to VGPR
// lanes and reads each back with v_readlane before every use, which costs
// VALU issue slots, s_nop hazards and VGPRs (measured: 239 v_readlane among
// 632 instructions of the fused step's stencil part).
//
// kernarg<T>(offset) returns a pointer into the kernel argument segment whose
// value the compiler cannot see through, so a member read through it is a
// scalar load (a scalar-cache hit) at the point of use, and nothing stays
// live across the loop.  Explicit kernel arguments start at offset 0 of the
// segment, so a kernel that takes one struct argument reads it through
// kernarg<Struct>(0) and its members at their C++ offsets.
#pragma once

#include <hip/hip_runtime.h>

namespace filt {

template <class T>
using KArg = const __attribute__((address_space(4))) T;

template <class T>
__device__ __forceinline__ KArg<T>* kernarg(unsigned long offset)
{
    KArg<T>* q = (KArg<T>*)((KArg<char>*)__builtin_amdgcn_kernarg_segment_ptr() + offset);
    asm volatile("" : "+s"(q));
    return q;
}

}  // namespace filt
