"""Writes reference_kats.json: the known-answer vectors that the reference's
own unit tests hold for the MH hot path, plus the Philox4x32-10 known-answer
vectors of the Random123 distribution (kat_vectors, SC'11).

Each entry cites the reference test it is taken from (paths relative to the
GLPhase source tree).  The values are data, restated from those assertions.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

kats = {
    "philox4x32_10": {
        "source": "Random123 kat_vectors (Salmon et al., SC'11)",
        "cases": [
            {"ctr": [0, 0, 0, 0], "key": [0, 0],
             "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
            {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
             "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
            {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
             "key": [0xa4093822, 0x299f31d0],
             "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
        ],
    },
    "unpack_with_codebook": {
        "source": "src/testing/cuda_tests.cpp:66-96",
        "bits": 4,
        "codebook": "cb[i] = (i/10, 2i/100) for i in 0..15, as float",
        "cases": [
            {"word": 1, "slot": 7, "gl": ["0.1", "0.02", "0.88"]},
        ] + [{"word": 19088743, "slot": i, "gl_index": i} for i in range(8)],
    },
    "fill_emit_ones": {
        "source": "src/testing/cuda_tests.cpp:98-116",
        "pc": "all ones",
        "expect": "emit[i] == GL0 + 2*GL1 + GL2 (EXPECT_FLOAT_EQ, 4 ulp)",
    },
    "mutation_matrix_layout": {
        "source": "src/testing/cuda_tests.cpp:46-63 and hmmLike.cu:341-344",
        "pc": "pc[i][j] = i + 4j; device copy h[i + 4j] = pc[i][j]",
    },
    "pack_nonvq": {
        "source": "src/testing/glPack_unittests.cpp:27-117",
        "one_site": {
            "gls": [[0, 1.0 / 16, 15.0 / 16], [2.0 / 16, 3.0 / 16, 11.0 / 16],
                    [4.0 / 16, 5.0 / 16, 7.0 / 16], [0, 0, 1], [0, 0, 1]],
            "stride": 3, "first_bytes": [0, 18, 52], "second_call_words": 2,
        },
        "two_sites": {
            "gls": [[[0, 1.0 / 16, 15.0 / 16], [2.0 / 16, 3.0 / 16, 11.0 / 16]],
                    [[4.0 / 16, 5.0 / 16, 7.0 / 16], [0.9 / 16, 3.5 / 16, 11.6 / 16]],
                    [[0, 0, 1], [0.5, 0.5, 0]]],
            "stride": 2,
            "bytes": {"s0_site0": 0, "s1_site0": 52, "s0_site1": 18, "s1_site1": 3,
                      "s2_site0": 0, "s2_site1": 119},
        },
    },
    "vq_find_points": {
        "source": "src/testing/glPack_unittests.cpp:128-163,252-290",
        "gls": [[0, 1.0 / 16, 15.0 / 16], [2.0 / 16, 3.0 / 16, 11.0 / 16],
                [4.0 / 16, 5.0 / 16, 7.0 / 16], [0.9 / 16, 3.5 / 16, 11.6 / 16],
                [0, 0, 1], [0.5, 0.5, 0]],
        "codebook_size": 16,
        "expect": "each point has a codeword within 0.01 in RR and Het, and FindGLCode returns "
                  "the first such codeword",
    },
    "euclidian_dist": {
        "source": "src/testing/glPack_unittests.cpp:119-126",
        "point": [0.5, 1.0], "a": [0, 0], "b": [2, 2], "a_closer": True,
    },
    "mh_sampler": {
        "source": "src/testing/other_unittests.cpp:256-283 and MHSampler.hpp:82-96",
        "rule": "accept when prop >= curr, else when u < exp((prop - curr) * pen)",
        "curr": -100.0, "better_always_accepted": True, "much_worse": -200.0,
    },
    "planted_parents": {
        "source": "src/testing/cuda_tests.cpp:118-656",
        "num_sites": 1024, "num_haps": 12, "num_cycles": 100, "high_exp": 100.0,
        "panel_words_all_ones": {"5": [0, 5], "6": [0, 5], "7": [1, 7], "8": [1, 7],
                                 "10": [2], "11": [3]},
        "alt_words": {"0": [0, 5, 2], "1": [1, 7, 3], "2": [0, 3, 5], "3": [1, 2, 7]},
        "expect": {"0": {"pair": [5, 6], "other": 10}, "1": {"pair": [7, 8], "other": 11},
                   "2": {"pair": [5, 6], "other": 11}, "3": {"pair": [7, 8], "other": 10}},
    },
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote reference_kats.json")
