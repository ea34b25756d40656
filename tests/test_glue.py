"""The multi-object restatement of diffrp's VAO and stencil-masked G-buffer
collect (CPU): one object reduces to the single-object collect, and two
objects give each object's single-object values on its own pixels."""
import numpy as np
import torch

import oracle
from oracle import glue
from tests.scenes import random_soup


def test_multi_object_collect_reduces_to_single_object_and_partitions(oracle_lib):
    pos, tri = random_soup(30, seed=60, spread=0.9)
    H, W = 40, 48
    rast = torch.from_numpy(oracle.rasterize(pos.numpy(), tri.numpy(), (H, W)))
    vals = torch.rand(len(pos), 3, generator=torch.Generator().manual_seed(61))
    one = torch.tensor([0] + [1] * len(tri), dtype=torch.int32)
    want, _ = glue.gbuffer_collect_single(rast, tri, one, lambda it: it.interpolate(vals), [0.25, 0.5, 0.75])
    got = glue.gbuffer_collect_layer(rast, tri, one, 1, lambda pi, it: it.interpolate(vals), [0.25, 0.5, 0.75])
    assert torch.equal(got, want)
    # two owners: triangles alternate between objects 1 and 2
    two = torch.tensor([0] + [1 + (t % 2) for t in range(len(tri))], dtype=torch.int32)
    scale = torch.tensor([1.0, 2.0, 3.0])
    got2 = glue.gbuffer_collect_layer(rast, tri, two, 2, lambda pi, it: it.interpolate(vals) * scale[pi],
                                      [0.25, 0.5, 0.75])
    owner = two[glue.float_to_triidx(rast[..., 3]).long()]
    for k in (1, 2):
        m = owner == k
        assert m.any()
        assert torch.equal(got2[m], want[m] * scale[k - 1])
    assert torch.equal(got2[owner == 0], want[owner == 0])
    # an object returning None keeps the default on its pixels
    got3 = glue.gbuffer_collect_layer(rast, tri, two, 2, lambda pi, it: None if pi == 0 else it.interpolate(vals),
                                      [0.25, 0.5, 0.75])
    assert torch.equal(got3[owner == 1], torch.tensor([0.25, 0.5, 0.75]).expand(int((owner == 1).sum()), 3))


def test_vertex_array_object_offsets_and_stencils():
    objs = []
    for k, (nv, nf) in enumerate([(5, 3), (4, 2), (6, 4)]):
        g = torch.Generator().manual_seed(70 + k)
        M = torch.eye(4)
        M[:3, 3] = torch.tensor([k, -k, 0.5 * k])
        objs.append({'verts': torch.rand(nv, 3, generator=g), 'tris': torch.randint(0, nv, (nf, 3), generator=g),
                     'M': M, 'color': torch.rand(nv, 4, generator=g)})
    vao = glue.vertex_array_object(objs)
    assert vao['stencils'].tolist() == [0, 1, 1, 1, 2, 2, 3, 3, 3, 3]
    assert torch.equal(vao['tris'][3:5], (objs[1]['tris'] + 5).int())
    assert torch.equal(vao['tris'][5:], (objs[2]['tris'] + 9).int())
    np.testing.assert_allclose(vao['world_pos'][9:].numpy(), (objs[2]['verts'] + objs[2]['M'][:3, 3]).numpy(),
                               rtol=1e-6)
    assert vao['color'].shape == (15, 4)
