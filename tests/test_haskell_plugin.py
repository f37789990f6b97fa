"""The Haskell plugin (haskell/src) against the C ABI it binds (include/afft.h), on the CPU.

GHC and the Accelerate packages are not available here, so the plugin is checked by reading it:
every `foreign import ccall` must name a function that include/afft.h declares, with the same
number of arguments, matching argument types and a matching return type; every libafft binding
the plugin modules use must be exported by the FFI module; every HIP.* name the front-end patch
and the Adjoint module use must be exported by the plugin; and the patches must apply to the
reference's FFT.hs and accelerate-fft.cabal when the reference tree is present."""
import glob
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HS = os.path.join(ROOT, "haskell")
HEADER = os.path.join(ROOT, "include", "afft.h")
PLUGIN = os.path.join(HS, "src", "Data", "Array", "Accelerate", "Math", "FFT", "LLVM", "HIP.hs")
FFI = os.path.join(HS, "src", "Data", "Array", "Accelerate", "Math", "FFT", "LLVM", "HIP", "FFI.hs")
ADJOINT = os.path.join(HS, "src", "Data", "Array", "Accelerate", "Math", "FFT", "Adjoint.hs")
REFERENCE = "/root/reference"

# Haskell FFI type -> C type (const dropped, pointers written "T *")
HS_TO_C = {
    "CInt": "int",
    "Int64": "int64_t",
    "CDouble": "double",
    "Ptr Int64": "int64_t *",
    "Ptr ()": "void *",
    "CString": "char *",
}


def c_prototypes():
    """name -> (return type, [argument types]) for every function include/afft.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    protos = {}
    for m in re.finditer(r"(const\s+char\s*\*\s*|int\s+)(afft_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        args = []
        for p in params.split(","):
            p = " ".join(p.replace("const", " ").split())
            if p in ("", "void"):
                continue
            stars = p.count("*")
            base = p.replace("*", " ").split()
            args.append(base[0] + (" " + "*" * stars if stars else ""))
        protos[name] = (norm_c(ret), args)
    return protos


def norm_c(t):
    t = " ".join(t.replace("const", " ").split())
    stars = t.count("*")
    base = t.replace("*", " ").split()[0]
    return base + (" " + "*" * stars if stars else "")


def hs_imports(path):
    """(C name, Haskell name, [argument types], return type) of every foreign import in a file."""
    text = open(path).read()
    out = []
    for m in re.finditer(r'foreign\s+import\s+ccall\s+(?:safe\s+|unsafe\s+)?"(\w+)"\s+(\w+)\s*::\s*(.*?)(?=\n\s*\n|\n--|\Z)',
                         text, flags=re.S):
        cname, hname, sig = m.group(1), m.group(2), " ".join(m.group(3).split())
        parts = [p.strip() for p in sig.split("->")]
        ret = parts[-1]
        assert ret.startswith("IO "), (hname, ret)
        out.append((cname, hname, parts[:-1], ret[3:].strip()))
    return out


def all_hs_files():
    return sorted(glob.glob(os.path.join(HS, "src", "**", "*.hs"), recursive=True))


def test_plugin_files_present():
    for f in (PLUGIN, FFI, ADJOINT,
              os.path.join(os.path.dirname(FFI), "Base.hs"),
              os.path.join(HS, "patches", "FFT.hs.patch"),
              os.path.join(HS, "patches", "accelerate-fft.cabal.patch")):
        assert os.path.exists(f), f


def test_every_foreign_import_matches_the_header():
    protos = c_prototypes()
    assert "afft_fft" in protos and "afft_spectral_conv_vjp" in protos
    seen = 0
    for path in all_hs_files():
        for cname, hname, args, ret in hs_imports(path):
            seen += 1
            assert cname in protos, f"{hname}: {cname} is not declared in afft.h"
            cret, cargs = protos[cname]
            assert len(args) == len(cargs), f"{cname}: {len(args)} arguments in Haskell, {len(cargs)} in C"
            for i, (h, c) in enumerate(zip(args, cargs)):
                assert h in HS_TO_C, f"{cname} argument {i}: unmapped Haskell type {h}"
                assert HS_TO_C[h] == c, f"{cname} argument {i}: {h} does not match C {c}"
            assert HS_TO_C.get(ret) == cret, f"{cname}: returns {ret}, C returns {cret}"
    assert seen >= 10


def test_front_end_bindings_cover_the_plugin_contract():
    names = {c for c, _, _, _ in hs_imports(FFI)}
    for need in ("afft_fft", "afft_fft1D", "afft_fft2D", "afft_fft3D", "afft_fft_vjp",
                 "afft_spectral_conv_fwd", "afft_spectral_conv_vjp", "afft_strerror"):
        assert need in names, need


def export_list(path):
    text = open(path).read()
    m = re.search(r"^module\s+[\w.]+\s*\((.*?)\)\s*where", text, flags=re.S | re.M)
    assert m, path
    body = re.sub(r"--[^\n]*", " ", m.group(1))
    return {w.strip() for w in re.split(r"[,\s]+", body) if w.strip()}


def test_plugin_modules_are_consistent():
    ffi_exports = export_list(FFI)
    ffi_defined = {h for _, h, _, _ in hs_imports(FFI)}
    assert {n for n in ffi_exports if n.startswith("afft_")} == ffi_defined
    # every libafft binding the plugin and its helpers use is exported by the FFI module
    for path in (PLUGIN, os.path.join(os.path.dirname(FFI), "Base.hs")):
        text = re.sub(r"--[^\n]*", " ", open(path).read())
        used = set(re.findall(r"\b(afft_\w+)\b", text))
        assert used <= ffi_exports, (path, used - ffi_exports)
    # every HIP.<name> the front-end patch and the Adjoint module use is exported by the plugin
    plugin_exports = export_list(PLUGIN)
    for path in (os.path.join(HS, "patches", "FFT.hs.patch"), ADJOINT):
        used = set(re.findall(r"\bHIP\.(\w+)", open(path).read()))
        used.discard("hs")  # the module name in an import line
        assert used and used <= plugin_exports, (path, used - plugin_exports)
    # every exported plugin function is defined with a type signature
    text = open(PLUGIN).read()
    for name in plugin_exports:
        assert re.search(rf"^{name}\s*::", text, flags=re.M), name


@pytest.mark.skipif(not os.path.isdir(REFERENCE) or shutil.which("patch") is None,
                    reason="the reference tree (or patch) is not present")
def test_patches_apply_to_the_reference(tmp_path):
    shutil.copytree(os.path.join(REFERENCE, "src"), tmp_path / "src")
    shutil.copy(os.path.join(REFERENCE, "accelerate-fft.cabal"), tmp_path / "accelerate-fft.cabal")
    for patch in ("FFT.hs.patch", "accelerate-fft.cabal.patch"):
        r = subprocess.run(["patch", "-p1", "--dry-run", "-i", os.path.join(HS, "patches", patch)],
                           cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
