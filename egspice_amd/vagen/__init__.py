"""vagen: the lane engine's Verilog-A compiler.

It reads a compact model's Verilog-A source (preprocessor, parser), splits
the analog block into the bias-independent setup (once per instance and
lane) and the bias-dependent evaluation (every Newton iteration), keeps
only what reaches a contribution, and emits C whose evaluation carries
sparse forward-mode derivatives with respect to the device's unknowns.
The reference compiles its models with VerA (an un-vendored dependency);
this compiler is the engine's own."""
from . import pp, parse, ir, compile as _compile, emit


def compile_file(path, prefix, defines=None):
    text = pp.preprocess(path, defines)
    mod = parse.parse(text)
    model = ir.Model(mod)
    c = _compile.Compiler(model, prefix)
    return emit.generate(c)
