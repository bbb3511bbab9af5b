"""Closed-form known-answer cases restated from TAMM's own tests.

Each case: index spaces, tensors with constant fills, one contraction, and the exact
expected value of every element of the output.  Sources:
  tests/tamm/Test_Ops.cpp:1400-1560  (three/four-dimensional mult ops, tilesize 1)
  tests/tamm/Test_Ops.cpp:1790-1880  (MultOp with RHS reduction labels)
  tests/tamm/Test_Mult_Ops.cpp:161-206 (4-D contraction, N=50 tile 20, norm 2.625e8)
Labels are single characters; 'S' denotes a scalar (0-mode) tensor.
"""

# (name, n, tile, clabels, cval, alpha, alabels, aval, blabels, bval, expected)
KATS = [
    ("mult_3x3x0", 10, 1, "ijk", 2.0, 6.9, "ijk", 3.0, "", 4.0, 2 + 6.9 * 3 * 4),
    ("mult_3x0x3", 10, 1, "ijk", 2.0, 1.7, "", 4.0, "ijk", 3.0, 2 + 1.7 * 3 * 4),
    ("mult_3x2x1", 10, 1, "ijk", 2.0, 1.7, "ij", 3.0, "k", 4.0, 2 + 1.7 * 3 * 4),
    ("mult_3x3x3_batch", 10, 1, "ijk", 2.0, 1.7, "jli", 3.0, "lik", 4.0, 2 + 1.7 * 3 * 4 * 10),
    ("mult_4x4x0", 10, 1, "ijkl", 2.0, 6.9, "ijkl", 3.0, "", 4.0, 2 + 6.9 * 3 * 4),
    ("mult_4x2x2", 10, 1, "ijkl", 2.0, 1.7, "ij", 3.0, "kl", 4.0, 2 + 1.7 * 3 * 4),
    ("mult_4x4x4_batch", 10, 1, "ijkl", 2.0, 1.7, "jlkm", 3.0, "likm", 4.0, 2 + 1.7 * 3 * 4 * 10),
    ("mult_2x2x2", 10, 1, "ij", 4.0, -3.1, "ik", 42.0, "jk", 5.0, 4 - 3.1 * 42 * 5 * 10),
    ("rhs_reduce_a", 4, 1, "", 4.0, -3.1, "i", 3.0, "", 5.0, 4.0 - 3.1 * 4 * 3 * 5),
    ("rhs_reduce_b", 4, 1, "", 4.0, -3.1, "", 3.0, "j", 5.0, 4.0 - 3.1 * 4 * 3 * 5),
    ("rhs_reduce_ab", 4, 1, "", 4.0, -3.2, "i", 3.0, "j", 5.0, 4.0 - 3.2 * 4 * 4 * 3 * 5),
    # Test_Ops.cpp:1639-1805, "Two-dimensional ops part I": T1 = 4, T2 = 42, T3 = 5.  T() takes
    # the positional default labels (labeled_tensor.hpp:213-216), so T1() and T2() share them;
    # "T1() -= x" is "T1() += -x".
    ("mult_2x2x0", 10, 1, "ij", 4.0, -3.1, "ij", 42.0, "", 5.0, 4 - 3.1 * 42 * 5),
    ("mult_2x2x0_minus", 10, 1, "ij", 4.0, 3.1, "ij", 42.0, "", 5.0, 4 + 3.1 * 42 * 5),
    ("mult_0x2x2", 10, 1, "", 4.0, -3.1, "ij", 42.0, "ij", 5.0, 4 - 3.1 * 10 * 10 * 42 * 5),
    ("mult_2x1x1", 10, 1, "ij", 4.0, -3.1, "i", 42.0, "j", 5.0, 4 - 3.1 * 42 * 5),
    ("mult_2x2x1_batch_i", 10, 1, "ij", 4.0, -3.1, "ij", 42.0, "i", 5.0, 4 - 3.1 * 42 * 5),
    ("mult_2x2x1_batch_i_perm", 10, 1, "ij", 4.0, -3.1, "ji", 42.0, "i", 5.0, 4 - 3.1 * 42 * 5),
    ("mult_2x2x1_batch_j", 10, 1, "ij", 4.0, -3.1, "ij", 42.0, "j", 5.0, 4 - 3.1 * 42 * 5),
    ("mult_2x1x2_batch_j", 10, 1, "ij", 4.0, -3.1, "j", 5.0, "ji", 42.0, 4 - 3.1 * 42 * 5),
    ("mult_2x2x2_hadamard", 10, 1, "ij", 4.0, -3.1, "ij", 42.0, "ij", 5.0, 4 - 3.1 * 42 * 5),
    # the same cases at a tile size that leaves ragged tiles
    ("mult_4x4x4_batch_t3", 10, 3, "ijkl", 2.0, 1.7, "jlkm", 3.0, "likm", 4.0, 2 + 1.7 * 3 * 4 * 10),
    ("mult_2x2x2_t4", 10, 4, "ij", 4.0, -3.1, "ik", 42.0, "jk", 5.0, 4 - 3.1 * 42 * 5 * 10),
]

# Test_Mult_Ops: C(j,i,k,l) += A(i,j,m,o) * B(m,o,k,l), A=21, B=2, N=50, tile 20.
MULT_OPS = dict(n=50, tile=20, c="jikl", a="ijmo", b="mokl", aval=21.0, bval=2.0, norm=2.625e8)


def label_ids(*strs):
    chars = []
    for s in strs:
        for ch in s:
            if ch not in chars:
                chars.append(ch)
    return {ch: -(i + 1) for i, ch in enumerate(chars)}
