// Sizes shared by the host engine and the neural-network kernels.
#pragma once
#include <stddef.h>

namespace ptmc {

constexpr int NNP_NP = 64;     // atoms per configuration (maximum)
constexpr int NNP_NT = 512;    // threads per workgroup
constexpr int NNP_MAXH = 32;   // hidden width (maximum)
constexpr int NNP_HS = NNP_MAXH + 1;  // padded hidden row stride
constexpr int NNP_GMAX = 8;    // functions per eta group (maximum)
constexpr int NNP_MAXF = 128;  // symmetry functions (maximum)
constexpr int NNP_PP = (NNP_NP * (NNP_NP - 1) / 2 + NNP_NT - 1) / NNP_NT;  // partner pairs per thread

// LDS layout in floats: X[3][NP] | G[nf][NP] | D[nf][NP] | W1[nf][h1] |
// H1[NP][HS] | H2[NP][HS] | rn2, fn, ro2, fo, eat, eatn [NP each]
inline size_t nnp_lds_floats(int nf, int h1)
{
    return (size_t)3 * NNP_NP + 2 * (size_t)nf * NNP_NP + (size_t)nf * h1 + 2 * (size_t)NNP_NP * NNP_HS +
           6 * NNP_NP;
}

}  // namespace ptmc
