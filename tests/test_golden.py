"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every fixture byte for byte, and each fixture's cross-checks (validators, brute
force, the reference's known answer 62) hold.  GPU: the HIP path reproduces every array of every fixture."""
import os

import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
NAMES = ["soup_1k", "soup_4k", "ico1_plane"]


def load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def same_bytes(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def _golden_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    f = load(name)
    g = _golden_module()
    new = g.generate(name)
    for k in f:
        assert same_bytes(f[k], new[k]), f"{name}: {k} changed"
    g.cross_check(name, f)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_path_reproduces_golden(ctx, name):
    print(f"[case] golden {name}", flush=True)
    f = load(name)
    sd, sdt, mp = (int(x) for x in f["params"])
    bvh2 = ob.PlocBuilder(ctx).build(sd, f["tris"], None, ob.SortPrecision.U64, sdt)
    codes, order = bvh2.sort_trace()
    assert same_bytes(codes, f["codes"]) and same_bytes(order, f["order"])
    assert same_bytes(bvh2.nodes, f["bvh2"]) and same_bytes(bvh2.primitive_indices, f["bvh2_prims"])
    cw = ob.bvh2_to_cwbvh(bvh2, mp, True, False)
    assert same_bytes(cw.nodes, f["cw"]) and same_bytes(cw.primitive_indices, f["cw_prims"])
    fused = ob.build_cwbvh_from_tris(ctx, f["tris"], ob.BvhBuildParams(ploc_search_distance=sd,
                                                                        search_depth_threshold=sdt,
                                                                        max_prims_per_leaf=mp))
    assert same_bytes(fused.nodes, f["cw"]) and same_bytes(fused.primitive_indices, f["cw_prims"])
    bvh_tris = cw.reorder_triangles(f["tris"])
    hits = cw.ray_traverse(f["rays"], bvh_tris)
    assert same_bytes(hits["primitive_id"], f["hits"]["primitive_id"])
    assert same_bytes(hits["t"], f["hits"]["t"])
    miss = cw.ray_traverse_miss(f["rays"], bvh_tris)
    assert same_bytes(miss.astype(np.uint8), f["miss"])
