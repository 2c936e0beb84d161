"""The oracle's ctypes structures keep the arrays they point to alive, so a
temporary (a converted G-buffer, a volume slice) can be passed straight into
an oracle call: a structure that only held the addresses let the arrays be
freed before the call read them."""
import ctypes as C
import gc

import numpy as np

import oracle_api as O


def _gbuffer(fill):
    return dict(diffuse=np.full((4, 5), fill, np.uint32), normal=np.zeros((4, 5, 4), np.uint16),
                specular=np.zeros((4, 5), np.uint32), emission=np.zeros((4, 5, 4), np.uint16),
                depth=np.full((4, 5), 0.5, np.float32))


def test_gbuf_struct_keeps_its_arrays():
    s = O.gbuf_struct(_gbuffer(7))
    gc.collect()
    churn = [_gbuffer(99) for _ in range(200)]  # reuse freed memory if the arrays were freed
    assert list((C.c_uint32 * 20).from_address(s.diffuse)) == [7] * 20
    assert list((C.c_float * 20).from_address(s.depth)) == [0.5] * 20
    del churn


def test_vol_and_tris_structs_keep_their_arrays():
    v = O.vol_struct(np.full((2, 3, 4), 5, np.uint32), 2, 1)
    t = O.tris_struct(np.full((1, 9), 1.5, np.float32), np.zeros((1, 9), np.float32), np.zeros(1, np.int32))
    gc.collect()
    churn = [np.full(64, 11, np.uint32) for _ in range(200)]
    assert list((C.c_uint32 * 24).from_address(v.texels)) == [5] * 24
    assert list((C.c_float * 9).from_address(t.pos)) == [1.5] * 9
    del churn
