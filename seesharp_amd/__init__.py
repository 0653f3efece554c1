"""seesharp_amd: an MI355X-native (gfx950) GPU path tracer that is a drop-in
`Integrator` for SeeSharp's scene model.

The product is libsspt.so (HIP kernels + C ABI, include/sspt.h); this package
is its Python host mirror of SeeSharp's object model.
"""
from .scene import (  # noqa: F401
    DiffuseEmitter,
    DiffuseMaterial,
    EnvironmentMap,
    FrameBuffer,
    GenericMaterial,
    GlossyEmitter,
    Mesh,
    PerspectiveCamera,
    Scene,
    TextureMono,
    TextureRgb,
    create_look_at,
)
from .integrator import (ClassicBidir, DebugVisualizer, Device, GpuBatch, GpuPathTracer, GpuScene,  # noqa: F401
                         Integrator, VertexCacheBidir, VertexCacheBidirLightTracer, VertexConnectionAndMerging)
