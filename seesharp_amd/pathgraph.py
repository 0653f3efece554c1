"""PathGraph (SeeSharp/Integrators/Util/PathGraph.cs): the tree of vertices a
replayed path visits, built from the node array sspt_replay_pixel returns.

Node classes carry the fields of the reference's node types: BSDFSampleNode
(ScatterWeight, SurvivalProbability, Emission, MISWeight), NextEventNode
(Emission, Pdf, BsdfTimesCosine, MISWeight, PrefixWeight, Contrib),
BackgroundNode (Emission, MISWeight). `PathGraph.ConvertToPLY` writes the
edges as an ascii PLY like the reference's (PathGraph.cs:163-250).
"""
from __future__ import annotations

import numpy as np

from . import abi


def _srgb_byte(c: float) -> int:
    """RgbColor.LinearToSrgb then (byte)(x * 255) (PathGraph.cs:173-180)."""
    c = max(0.0, min(1.0, c))
    s = 12.92 * c if c <= 0.0031308 else 1.055 * c ** (1 / 2.4) - 0.055
    return int(s * 255)


def _linear(r: int, g: int, b: int):
    """RgbColor.SrgbToLinear of 8-bit sRGB components."""
    def f(v):
        v = v / 255.0
        return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4
    return (f(r), f(g), f(b))


class PathGraphNode:
    VisualizerColor = (0.0, 0.0, 0.0)
    IsBackground = False

    def __init__(self, position):
        self.Position = np.asarray(position, np.float32)
        self.Ancestor: PathGraphNode | None = None
        self.Successors: list[PathGraphNode] = []

    def AddSuccessor(self, node: "PathGraphNode") -> "PathGraphNode":
        self.Successors.append(node)
        node.Ancestor = self
        return node

    def ComputeVisualizerColor(self):
        return self.VisualizerColor


class BSDFSampleNode(PathGraphNode):
    VisualizerColor = _linear(41, 107, 177)

    def __init__(self, position, scatter_weight, survival_probability, emission=(0.0, 0.0, 0.0), mis_weight=0.0):
        super().__init__(position)
        self.ScatterWeight = np.asarray(scatter_weight, np.float32)
        self.SurvivalProbability = float(survival_probability)
        self.Emission = np.asarray(emission, np.float32)
        self.MISWeight = float(mis_weight)


class NextEventNode(PathGraphNode):
    VisualizerColor = _linear(170, 231, 232)

    def __init__(self, position, emission, pdf, bsdf_cos, mis_weight, prefix_weight, background: bool):
        super().__init__(position)
        self.Emission = np.asarray(emission, np.float32)
        self.Pdf = float(pdf)
        self.BsdfTimesCosine = np.asarray(bsdf_cos, np.float32)
        self.MISWeight = float(mis_weight)
        self.PrefixWeight = np.asarray(prefix_weight, np.float32)
        self.IsBackground = background

    @property
    def Contrib(self):
        """PrefixWeight * Emission / Pdf * BsdfTimesCosine (PathGraph.cs:65)."""
        return self.PrefixWeight * self.Emission / np.float32(self.Pdf) * self.BsdfTimesCosine


class BackgroundNode(PathGraphNode):
    VisualizerColor = _linear(170, 231, 232)
    IsBackground = True

    def __init__(self, position, emission, mis_weight):
        super().__init__(position)
        self.Emission = np.asarray(emission, np.float32)
        self.MISWeight = float(mis_weight)


class PathGraph:
    def __init__(self):
        self.Roots: list[PathGraphNode] = []

    @staticmethod
    def from_nodes(nodes) -> "PathGraph":
        """Builds the tree from sspt_path_node records (ancestors come first)."""
        g = PathGraph()
        made: list[PathGraphNode] = []
        for n in nodes:
            pos = tuple(n.position)
            if n.kind == abi.NODE_CAMERA:
                x = PathGraphNode(pos)
            elif n.kind == abi.NODE_BSDF_SAMPLE:
                x = BSDFSampleNode(pos, tuple(n.scatter_weight), n.survival_probability, tuple(n.emission),
                                   n.mis_weight)
            elif n.kind == abi.NODE_NEXT_EVENT:
                x = NextEventNode(pos, tuple(n.emission), n.pdf, tuple(n.bsdf_cos), n.mis_weight,
                                  tuple(n.prefix_weight), bool(n.flags & abi.NODE_TO_BACKGROUND))
            elif n.kind == abi.NODE_BACKGROUND:
                x = BackgroundNode(pos, tuple(n.emission), n.mis_weight)
            else:
                raise ValueError(f"unknown node kind {n.kind}")
            if n.ancestor < 0:
                g.Roots.append(x)
            else:
                if not 0 <= n.ancestor < len(made):
                    raise ValueError("node ancestor out of order")
                made[n.ancestor].AddSuccessor(x)
            made.append(x)
        return g

    def nodes(self):
        """Every node, depth first in insertion order."""
        out = []

        def walk(n):
            out.append(n)
            for s in n.Successors:
                walk(s)
        for r in self.Roots:
            walk(r)
        return out

    def ConvertToPLY(self) -> str:
        """Ascii PLY of the graph's edges, two coloured vertices per edge (PathGraph.cs:226-250)."""
        verts, edges = [], []

        def add(node):
            for s in node.Successors:
                r, g, b = (_srgb_byte(c) for c in s.ComputeVisualizerColor())
                verts.append((node.Position, r, g, b))
                verts.append((s.Position, r, g, b))
                edges.append((len(verts) - 2, len(verts) - 1))
                add(s)
        for root in self.Roots:
            add(root)
        head = ("ply\nformat ascii 1.0\n"
                f"element vertex {len(verts)}\nproperty float x\nproperty float y\nproperty float z\n"
                "property uchar red\nproperty uchar green\nproperty uchar blue\n"
                f"element edge {len(edges)}\nproperty int vertex1\nproperty int vertex2\nend_header\n")
        vs = "\n".join(f"{p[0]} {p[1]} {p[2]} {r} {g} {b}" for p, r, g, b in verts)
        es = "\n".join(f"{a} {b}" for a, b in edges)
        return head + vs + "\n" + es
