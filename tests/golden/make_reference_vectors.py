"""Writes reference_vectors.json: the known-answer vectors that the reference's own
tests hold for the TLR GEMM path, transcribed as data (row-major, as written there).

Sources (ecrc/hcorepp):
  tests/api/TestGemm.cpp                      sections "Gemm test 1" .. "Gemm Test 10"
  tests/operators/TestCompressedTile.cpp:124-260   "Compressed Tile Gemm"
  tests/kernels/TestKernels.cpp               ProcessVpointer :157-192, CalculateNewRank :194-237,
                                              CalculateUVptr :239-265, CalculateVTnew :267-329,
                                              LaCpy :383-468, geqrf :470-511, Laset :513-604,
                                              Potrf :1070-1124, Trsm :1126-1183, CrossProduct (syrk) :1185-1248
The reference validates products with a 1 % relative tolerance
(tests/helpers/src/testHelpers.cpp:83-86); exact integer kernels are compared exactly.
"""
import json
import os

GEMM = [
    dict(name="test1_DDD", src="tests/api/TestGemm.cpp:27-84", kind="DDD",
         A=[[1, 2, 3], [4, 5, 6], [7, 8, 9]], B=[[2, 4, 6], [8, 10, 12], [14, 16, 18]],
         out=[[60, 72, 84], [132, 162, 192], [204, 252, 300]]),
    dict(name="test2_DDD", src="tests/api/TestGemm.cpp:86-144", kind="DDD",
         A=[[5], [10], [15]], B=[[2, 4, 6]], out=[[10, 20, 30], [20, 40, 60], [30, 60, 90]]),
    dict(name="test3_DDD", src="tests/api/TestGemm.cpp:146-204", kind="DDD",
         A=[[10, 11, 12]], B=[[2, 4], [8, 10], [14, 16]], out=[[276, 342]]),
    dict(name="test4_CDD", src="tests/api/TestGemm.cpp:206-287", kind="CDD",
         Au=[[1], [4], [7]], Av=[[10, 11, 12]], B=[[2, 4], [8, 10], [14, 16]],
         out=[[276, 342], [1104, 1368], [1932, 2394]]),
    dict(name="test5_DCD", src="tests/api/TestGemm.cpp:289-368", kind="DCD",
         A=[[2, 1, 4], [8, 5, 10], [14, 10, 16]], Bu=[[1], [4], [7]], Bv=[[10, 11, 12]],
         out=[[340, 374, 408], [980, 1078, 1176], [1660, 1826, 1992]]),
    dict(name="test6_CCD", src="tests/api/TestGemm.cpp:370-466", kind="CCD",
         Au=[[2, 4], [8, 10], [14, 16]], Av=[[1, 5, 6, 7], [3, 4, 8, 9]],
         Bu=[[2, 10, 18], [4, 12, 20], [6, 14, 22], [8, 16, 24]], Bv=[[5, 25], [10, 30], [15, 35]],
         out=[[66760, 178840], [195400, 523480], [324040, 868120]]),
    dict(name="test7_CDC", src="tests/api/TestGemm.cpp:468-603", kind="CDC", c_rank=1, acc="eps",
         Au=[[2, 8], [4, 10], [6, 12]], Av=[[5, 15, 25], [10, 20, 30]], B=[[1, 9], [5, 10], [7, 12]],
         out=[[3070, 6190], [4220, 8480], [5370, 10770]]),
    dict(name="test8_DCC", src="tests/api/TestGemm.cpp:605-746", kind="DCC", c_rank=1, acc="eps",
         A=[[2, 8, 14, 20], [4, 10, 16, 22], [6, 12, 18, 24]], Bu=[[1, 9], [3, 11], [5, 13], [7, 15]],
         Bv=[[5, 15], [10, 20]], out=[[7060, 15300], [8180, 17700], [9300, 20100]]),
    dict(name="test9_CCC", src="tests/api/TestGemm.cpp:748-915", kind="CCC", c_rank=1, acc="eps",
         Au=[[1, 10, 20], [2, 11, 21], [3, 12, 22], [4, 13, 23], [5, 14, 24]],
         Av=[[2, 8, 14, 20], [4, 10, 16, 22], [6, 12, 18, 24]], Bu=[[1, 9], [3, 11], [5, 13], [7, 15]],
         Bv=[[5, 15], [10, 20]],
         out=[[274860, 594300], [299400, 647400], [323940, 700500], [348480, 753600], [373020, 806700]]),
    dict(name="test10_DDC", src="tests/api/TestGemm.cpp:917-1052", kind="DDC", c_rank=4, acc="eps",
         A=[[1, 10, 20], [2, 11, 21], [3, 12, 22], [4, 13, 23], [5, 14, 24]],
         B=[[2, 8, 14, 20], [4, 10, 16, 22], [6, 12, 18, 24]],
         out=[[162, 348, 534, 720], [174, 378, 582, 786], [186, 408, 630, 852], [198, 438, 678, 918],
              [210, 468, 726, 984]]),
]

# TestCompressedTile "Compressed Tile Gemm": CompressedTile::Gemm called directly with a dense
# 3x3 A (row-major literal {{1,4,7},{2,5,8},{3,6,9}}), a dense 3x2 "B" (arank = 3), zero C of rank 2.
COMPRESSED_DIRECT = dict(name="compressed_tile_gemm", src="tests/operators/TestCompressedTile.cpp:124-260",
                         A=[[1, 4, 7], [2, 5, 8], [3, 6, 9]], B=[[2, 8], [4, 10], [6, 12]], c_rank=2,
                         acc="eps", out=[[60, 132], [72, 162], [84, 192]])

KERNELS = dict(
    process_vpointer=dict(src="tests/kernels/TestKernels.cpp:157-192", n=2, crank=2, arank=3, beta=1.0,
                          cv=[1, 2, 3, 4], b=[5, 6, 7, 8, 9, 10], expected=[1, 3, 2, 4, 5, 8, 6, 9, 7, 10]),
    new_rank_abs=dict(src="tests/kernels/TestKernels.cpp:194-214", acc=10.0, sigma=[11, 12, 10, 9, 8],
                      truncated=False, expected=3),
    new_rank_rel=dict(src="tests/kernels/TestKernels.cpp:216-237", acc=10.0, sigma=[2, 22, 20, 19, 18],
                      truncated=True, expected=3),
    uvptr=dict(src="tests/kernels/TestKernels.cpp:239-265", vm=2, rank=3, vnew=[1, 2, 3, 4, 5, 6],
               expected=[1, 3, 5, 2, 4, 6]),
    vtnew_ungqr=dict(src="tests/kernels/TestKernels.cpp:301-329", rank=3, min_vm_vn=3, size_s=3,
                     sigma=[2, 3, 0], vt=[1, 2, 3, 4, 5, 6, 7, 8, 9, 1, 2, 3],
                     expected=[2, 6, 0, 8, 15, 0, 14, 24, 0, 1, 2, 3]),
    lacpy_upper=dict(src="tests/kernels/TestKernels.cpp:415-442", m=4, n=4, a=list(range(1, 17)),
                     expected=[1, 0, 0, 0, 5, 6, 0, 0, 9, 10, 11, 0, 13, 14, 15, 16]),
    geqrf=dict(src="tests/kernels/TestKernels.cpp:470-511", m=6, n=2,
               a_rowmajor_T=[[0, 2, 2, 0, 2, 2], [2, -1, -1, 1.5, -1, -1]],
               expected_rowmajor_T=[[-4, 0.5, 0.5, 0, 0.5, 0.5], [2, 2.5, 0.285714, -0.428571, 0.285714, 0.285714]],
               tau=[1, 1.4]),
    laset_lower=dict(src="tests/kernels/TestKernels.cpp:578-604", m=4, n=4, off=1.0, diag=2.0,
                     expected=[2, 1, 1, 1, 0, 2, 1, 1, 0, 0, 2, 1, 0, 0, 0, 2]),
    svd_3x3=dict(src="tests/kernels/TestKernels.cpp:696-790", a=[[1, 2, 3], [4, 5, 6], [7, 8, 9]]),
    # Cholesky tile kernels (row-major literals as written there)
    potrf_lower=dict(src="tests/kernels/TestKernels.cpp:1070-1124", a=[[25, 15, 5], [15, 13, 11], [5, 11, 21]],
                     expected=[[5, 0, 0], [3, 2, 0], [1, 4, 2]]),
    trsm_left_upper=dict(src="tests/kernels/TestKernels.cpp:1126-1183", a=[[1, 2, 3], [0, 1, 1], [0, 0, 2]],
                         b=[[8], [4], [2]], expected=[[-1], [3], [1]]),
    syrk_lower=dict(src="tests/kernels/TestKernels.cpp:1185-1248",
                    a=[[3.12393, -1.16854, -0.304408, -2.15901], [-1.16854, 1.86968, 1.04094, 1.35925],
                       [-0.304408, 1.04094, 4.43374, 1.21072], [-2.15901, 1.35925, 1.21072, 5.57265]],
                    expected=[[15.878412787064, -1.16854, -0.304408, -2.15901],
                              [-9.08673783542, 7.7923056801, 1.04094, 1.35925],
                              [-6.13095182416, 8.56286609912, 22.300113620064, 1.21072],
                              [-20.73289403456, 13.8991634747, 14.18705411188, 39.0291556835]]),
)

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.json")
    with open(out, "w") as f:
        json.dump(dict(gemm=GEMM, compressed_direct=COMPRESSED_DIRECT, kernels=KERNELS), f, indent=1)
    print("wrote", out)
