"""The traversal node encoding (host side, seesharp_amd/csrc/bvh_build.cpp).

The traversal kernels read each 4-wide node as 80 bytes: the node origin
(float), a power-of-two scale per axis and every child plane as an fp16
offset, rounded outward. A box test on those planes is conservative only if
every decoded plane encloses the builder's float plane, so this checks, in
exact rational arithmetic, origin + h * 2^e <= lo and origin + h * 2^e >= hi
for every child plane of trees built from triangle soups (random sizes and
positions, far-from-origin clusters, flat and coincident triangles, leaves of
more than four triangles), that the outward rounding stays within a few fp16
steps (a wrong exponent or rounding direction fails one of the two bounds),
that only zero and normal fp16 values are stored, and that the slot kinds and
child words carry the float tree's children. The encoder and the host
builder are compiled from the library's sources with g++ and run on the CPU.
"""
import os
import struct
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "seesharp_amd", "csrc")

DRIVER = r"""
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include "bvh_build.h"
using namespace sspt;
int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    int32_t ntri = 0, max_leaf = 1;
    if (std::fread(&ntri, 4, 1, f) != 1 || std::fread(&max_leaf, 4, 1, f) != 1) return 3;
    std::vector<float> v(9 * (size_t)ntri);
    if (std::fread(v.data(), 4, v.size(), f) != v.size()) return 3;
    std::fclose(f);
    BvhOutput out;
    build_bvh(BvhInput{v.data(), ntri}, max_leaf, 4, &out);
    std::vector<uint32_t> words;
    std::vector<int> big;
    std::string err;
    if (!encode_traversal_nodes(out.nodes, 4, &words, &big, &err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 2;
    }
    FILE* g = std::fopen(argv[2], "wb");
    uint64_t nn = out.nodes.size() / 32, nb = big.size();
    std::fwrite(&nn, 8, 1, g);
    std::fwrite(out.nodes.data(), 4, out.nodes.size(), g);
    std::fwrite(words.data(), 4, words.size(), g);
    std::fwrite(&nb, 8, 1, g);
    std::fwrite(big.data(), 4, big.size(), g);
    std::fclose(g);
    return 0;
}
"""


@pytest.fixture(scope="module")
def encoder(tmp_path_factory):
    d = tmp_path_factory.mktemp("trav_nodes")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "encode"
    cmd = ["g++", "-std=c++17", "-O2", "-pthread", "-ffp-contract=off", f"-I{CSRC}", "-I/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", "-o", str(exe), str(src), os.path.join(CSRC, "bvh_build.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("compiling the encoder failed:\n" + r.stderr[-2000:])

    def run(verts, max_leaf=1):
        verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
        inp, out = d / "in.bin", d / "out.bin"
        with open(inp, "wb") as f:
            f.write(struct.pack("<ii", len(verts), max_leaf))
            f.write(verts.tobytes())
        subprocess.run([str(exe), str(inp), str(out)], check=True, capture_output=True)
        raw = out.read_bytes()
        nn = struct.unpack_from("<Q", raw, 0)[0]
        off = 8
        nodes = np.frombuffer(raw, np.float32, nn * 32, off).reshape(nn, 32)
        off += nn * 128
        words = np.frombuffer(raw, np.uint32, nn * 20, off).reshape(nn, 20)
        off += nn * 80
        nb = struct.unpack_from("<Q", raw, off)[0]
        big = np.frombuffer(raw, np.int32, nb, off + 8).reshape(-1, 2)
        return nodes, words, big

    return run


def f16_value(b):
    b = int(b)
    if b == 0:
        return Fraction(0)
    return Fraction(1024 + (b & 1023), 1024) * Fraction(2) ** ((b >> 10) - 15)


def counts_of(nodes):
    c = nodes[:, 28:32].view(np.int32).copy()
    c[:, 0] = (c[:, 0] << 8) >> 8
    return c


def check_tree(nodes, words, big, sample=None, seed=0):
    nn = len(nodes)
    counts = counts_of(nodes)
    child = nodes[:, 24:28].view(np.int32)
    # slot kinds and child words: every node
    meta = words[:, 3] >> 24
    leafk, innerk = meta & 15, meta >> 4
    cw = words[:, 16:20].view(np.int32)
    for k in range(4):
        c = counts[:, k]
        lk, ik = (leafk >> k) & 1, (innerk >> k) & 1
        absent, inner, small, large = c < 0, c == 0, (c > 0) & (c <= 4), c > 4
        assert np.all((lk == 0) & (ik == 0) | ~absent) and np.all(cw[absent, k] == -1)
        assert np.all(ik[inner] == 1) and np.all(lk[inner] == 0) and np.all(cw[inner, k] == child[inner, k])
        assert np.all(lk[small] == 1) and np.all(ik[small] == 0)
        expect = (child[small, k].astype(np.int64) | ((c[small].astype(np.int64) - 1) << 30)).astype(np.uint32)
        assert np.array_equal(cw[small, k].view(np.uint32), expect)
        assert np.all(lk[large] == 1) and np.all(ik[large] == 1)
        if np.any(large):
            b = cw[large, k]
            assert np.array_equal(big[b, 0], child[large, k]) and np.array_equal(big[b, 1], c[large])
    # fp16 offsets: only zero and normal values (no subnormals, infinities or NaNs)
    halves = np.concatenate([words[:, 4:16] & 0xFFFF, words[:, 4:16] >> 16], axis=1)
    exps = (halves >> 10) & 31
    assert np.all((halves == 0) | ((exps > 0) & (exps < 31))) and np.all(halves >> 15 == 0)
    # exact containment and tightness on a sample of nodes
    idx = np.arange(nn)
    if sample is not None and nn > sample:
        idx = np.random.default_rng(seed).choice(nn, sample, replace=False)
    checked = 0
    for n in idx:
        w = words[n]
        e = [int(np.int8(np.uint8((int(w[3]) >> (8 * a)) & 255))) for a in range(3)]
        for a in range(3):
            origin = Fraction(float(w[a:a + 1].view(np.float32)[0]))
            scale = Fraction(2) ** e[a]
            lo = nodes[n, 8 * a:8 * a + 4]
            hi = nodes[n, 8 * a + 4:8 * a + 8]
            valid = counts[n] >= 0
            ext = max(Fraction(float(hi[k])) for k in range(4) if valid[k]) - origin
            for k in range(4):
                if not valid[k]:
                    continue
                wl, wh = int(w[4 + 4 * a + k // 2]), int(w[6 + 4 * a + k // 2])
                hl = (wl >> (16 * (k & 1))) & 0xFFFF
                hh = (wh >> (16 * (k & 1))) & 0xFFFF
                dlo = origin + f16_value(hl) * scale
                dhi = origin + f16_value(hh) * scale
                flo, fhi = Fraction(float(lo[k])), Fraction(float(hi[k]))
                assert dlo <= flo, (n, a, k, "lower plane rounded inward")
                assert dhi >= fhi, (n, a, k, "upper plane rounded inward")
                # outward by at most two fp16 steps at the offset's magnitude (<= 2^-9 of the
                # node extent), or up to the smallest normal fp16 for tiny offsets
                slack = max(ext / 256, Fraction(2) ** -13 * scale)
                assert flo - dlo <= slack and dhi - fhi <= slack, (n, a, k, "rounded too far")
                checked += 1
    return checked


def soup(rng, n, centre, spread, size):
    c = centre + rng.uniform(-spread, spread, (n, 1, 3))
    return (c + rng.normal(0.0, 1.0, (n, 3, 3)) * size).reshape(n, 9)


def test_random_soup_planes_enclose_children(encoder):
    rng = np.random.default_rng(11)
    v = np.concatenate([soup(rng, 3000, 0.0, 5.0, 0.05), soup(rng, 1000, -3.0, 0.5, 0.002),
                        soup(rng, 500, 2.0, 50.0, 1.0)])
    nodes, words, big = encoder(v)
    assert len(big) == 0
    assert check_tree(nodes, words, big) > 10000


def test_far_from_origin_clusters(encoder):
    # small triangles a million units out: the node extents are tiny against the coordinates
    rng = np.random.default_rng(12)
    v = np.concatenate([soup(rng, 1500, 1.0e6, 1.0, 1e-3), soup(rng, 1500, -2.5e5, 0.01, 1e-5)])
    nodes, words, big = encoder(v)
    assert check_tree(nodes, words, big) > 5000


def test_flat_and_coincident_triangles(encoder):
    rng = np.random.default_rng(13)
    flat = soup(rng, 1200, 0.0, 2.0, 0.1).reshape(-1, 3, 3)
    flat[:, :, 1] = 0.75  # every triangle in the plane y = 0.75: zero extent on y
    same = np.tile(np.array([0.1, 0.2, 0.3, 0.4, 0.2, 0.3, 0.1, 0.6, 0.3], np.float32), (64, 1))
    nodes, words, big = encoder(np.concatenate([flat.reshape(-1, 9), same]))
    assert check_tree(nodes, words, big) > 3000


def test_leaves_of_more_than_four_triangles(encoder):
    rng = np.random.default_rng(14)
    nodes, words, big = encoder(soup(rng, 4000, 0.0, 3.0, 0.05), max_leaf=8)
    assert len(big) > 0 and np.all(big[:, 1] > 4) and np.all(big[:, 1] <= 8)
    assert check_tree(nodes, words, big, sample=1500) > 3000


def test_large_scene_planes_enclose_children(encoder):
    from seesharp_amd import scenes

    s = scenes.large_scene(16, 16, num_spheres=80, grid=96)
    v = np.concatenate([m.Vertices[m.Indices.reshape(-1, 3)].reshape(-1, 9) for m in s.Meshes])
    assert len(v) > 100000
    nodes, words, big = encoder(v)
    assert check_tree(nodes, words, big, sample=4000, seed=3) > 30000
