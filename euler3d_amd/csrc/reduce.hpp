// Deterministic sums on 64-lane wavefronts (gfx950).
//
// waveSum: the butterfly x += shfl_xor(x, 32), 16, 8, 4, 2, 1. At each step two partner lanes
// add the same two values (in either order, which IEEE addition does not distinguish), so every
// lane ends with the same bits, whatever the scheduling.
// groupSum<N>: each wave of an N-thread workgroup reduces its 64 values with waveSum, lane 0 puts
// the wave's total in lds[w], and after one barrier every thread adds the N / 64 totals in wave
// order. lds must hold N / 64 doubles that no other thread touches meanwhile; successive calls
// on the same lds need a barrier between them.
#pragma once

#include <hip/hip_runtime.h>

namespace e3d {

__device__ __forceinline__ double waveSum(double x) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
	return x;
}

template<int N>
__device__ __forceinline__ double groupSum(double x, double *lds) {
	static_assert(N % 64 == 0, "whole wavefronts");
	x = waveSum(x);
	if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
	__syncthreads();
	double s = lds[0];
#pragma unroll
	for (int w = 1; w < N / 64; w++) s += lds[w];
	return s;
}

}// namespace e3d
