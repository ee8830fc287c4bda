"""A small matrix-arithmetic program for the runtime-compilation tests: compiler-option and
external-header checks that fail on purpose, a __constant__ global and template kernels looked up
by lowered name (the behaviour SuperTest+/SuperTerrain+/STPTestRTC.cpp:72-292 checks)."""

HEADER_NAME = "MatrixOperatorVersion"
HEADER = '#pragma once\n#define MATRIX_OPERATOR_VERSION "4.2.1"\n'
NAMES = ["MatrixDimension", "transform<MatrixOp::Add>", "transform<MatrixOp::Sub>", "scale"]

SOURCE = r"""
#ifndef __HIPCC_RTC__
#error this program is compiled by hipRTC only
#endif
#ifndef STP_TEST_ENABLE
#error STP_TEST_ENABLE is not defined by the compiler
#endif
#ifdef STP_TEST_EXTERNAL_HEADER
#include <MatrixOperatorVersion>
#ifdef MATRIX_OPERATOR_VERSION
#pragma message("matrix operator version " MATRIX_OPERATOR_VERSION)
#error PASS
#else
#error MATRIX_OPERATOR_VERSION is not defined
#endif
#endif

enum class MatrixOp : unsigned char { Add = 0u, Sub = 1u };

__constant__ uint2 MatrixDimension[1];

__device__ inline bool element(unsigned int& index) {
	const unsigned int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	index = x + y * MatrixDimension[0].x;
	return x < MatrixDimension[0].x && y < MatrixDimension[0].y;
}

template <MatrixOp Op>
__global__ void transform(float* out, const float* a, const float* b);

template <>
__global__ void transform<MatrixOp::Add>(float* out, const float* a, const float* b) {
	unsigned int i;
	if (element(i)) {
		out[i] = a[i] + b[i];
	}
}

template <>
__global__ void transform<MatrixOp::Sub>(float* out, const float* a, const float* b) {
	unsigned int i;
	if (element(i)) {
		out[i] = a[i] - b[i];
	}
}

__global__ void scale(float* out, const float* in, float factor) {
	unsigned int i;
	if (element(i)) {
		out[i] = in[i] * factor;
	}
}
"""
