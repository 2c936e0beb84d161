"""GPU SpecularFilterPass (vct_specular_filter) against the oracle: both
filter methods, tonemap on and off, on random images whose size is not a
multiple of the 16-pixel tile; FP32 on both sides, equal to 1e-5."""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,tonemap", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_specular_filter_matches_oracle(method, tonemap):
    import torch
    import vk_voxel_cone_tracing_amd as V

    H, W = 61, 97
    rng = np.random.default_rng(10 + 2 * method + tonemap)
    diffuse = rng.uniform(0, 2, (H, W, 4)).astype(np.float32)
    spec = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    spec[20:40, 30:60, :3] += 3.0  # an edge for the bilateral weights
    d_t, s_t = torch.from_numpy(diffuse).cuda(), torch.from_numpy(spec).cuda()
    out = torch.empty_like(d_t)
    desc = V._abi.SpecularFilterDesc(2.2, 0.1, tonemap, method)
    V._abi.call("vct_specular_filter", None, C.c_void_p(d_t.data_ptr()), C.c_void_p(s_t.data_ptr()), W, H,
                C.byref(desc), C.c_void_p(out.data_ptr()))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = O.specular_filter(diffuse, spec, 2.2, 0.1, tonemap, method)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
