"""MI355X-native inter-analysis path of the Chicago/EVX3 encoder.

`chicago_amd.evxa` binds libevxa.so (HIP kernels for gfx950 behind the C ABI
in include/evxa.h); `chicago_amd.synth` builds seeded synthetic content;
`chicago_amd.build` compiles the library in-tree.
"""
__all__ = ["evxa", "synth", "build"]
