"""The C# binding shown in INTEGRATION.md declares exactly the functions of
include/matflat_hip.h, with the same parameter counts and types (CPU only:
the C# is parsed, not compiled -- there is no .NET toolchain here)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_decls():
    text = open(os.path.join(ROOT, "include", "matflat_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for ret, name, params in re.findall(r"^\s*(int|const char\*)\s+(mfh_\w+)\s*\(([^;]*?)\)\s*;", text, re.M | re.S):
        ps = [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]
        types = []
        for p in ps:
            ty = re.match(r"(.*?)\w+$", p).group(1)
            types.append(ty.replace("const", "").replace(" ", ""))
        out[name] = ("IntPtr" if ret.startswith("const char") else "int", types)
    return out


CS_OF_C = {"mfh_handle*": "IntPtr", "mfh_plan*": "IntPtr", "void*": "IntPtr", "mfh_handle**": "IntPtr*",
           "mfh_plan**": "IntPtr*", "void**": "IntPtr*", "int": "int", "int*": "int*", "int**": "int**",
           "double": "double", "double*": "double*", "double**": "double**", "float*": "float*", "int64_t": "long"}


def cs_decls():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = text[text.index("internal static unsafe partial class Native"):]
    block = block[:block.index("```")]
    out = {}
    for ret, name, params in re.findall(r"\[LibraryImport\(Lib\)\]\s*internal static partial (\w+) (mfh_\w+)\((.*?)\);",
                                        block, re.S):
        ps = [p.strip() for p in params.replace("\n", " ").split(",") if p.strip()]
        out[name] = (ret, [re.match(r"(.*?)\s*\w+$", p).group(1).replace(" ", "") for p in ps])
    return out


def test_every_header_function_is_bound_with_matching_types():
    h, cs = header_decls(), cs_decls()
    assert len(h) >= 40
    assert sorted(h) == sorted(cs), (sorted(set(h) ^ set(cs)))
    for name, (ret, types) in h.items():
        cret, ctypes_ = cs[name]
        assert cret == ret, name
        assert [CS_OF_C[t] for t in types] == ctypes_, name


def test_public_surface_names_the_reference_operations():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    for api in ("public static void Lu(", "public static void Cholesky(", "public static void SolveTriangular(",
                "public static void MulMatMat(", "public static void MulMatVec(", "public static void LuInverse(",
                "public static void SpdSolveBackward(", "public static Plan Lu(", "Complex** a"):
        assert api in text, api
