"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/hpcsort.h declares, and rejects bad arguments without a GPU."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from hpcsharp_amd import _native as N

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "hpcsort.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpcs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    assert declared_functions() == sorted(N.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libhpcsort.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    lib = N.lib()
    for f in declared_functions():
        assert getattr(lib, f) is not None


def test_csharp_binding_imports_only_declared_entry_points():
    """The C# binding given in INTEGRATION.md imports entry points by name:
    each must be one that include/hpcsort.h declares."""
    with open(os.path.join(REPO, "INTEGRATION.md")) as f:
        integ = f.read()
    with open(os.path.join(REPO, "include", "hpcsort.h")) as f:
        declared = set(re.findall(r"\b(hpcs_\w+)\s*\(", f.read()))
    imported = set(re.findall(r"partial (?:int|IntPtr) (hpcs_\w+)\(", integ))
    assert imported, "no LibraryImport declarations found"
    assert imported <= declared, sorted(imported - declared)


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_null_context_is_rejected_without_a_gpu():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libhpcsort.so not built")
    lib = N.lib()
    a = np.arange(4, dtype=np.uint32)
    rc = lib.hpcs_sort_radix_u32(None, a.ctypes.data_as(ctypes.c_void_p), 4)
    assert rc == N.HPCS_E_INVALID_ARG
    assert b"null" in lib.hpcs_last_error()
    assert lib.hpcs_version() >= 100
    info = N.PlanInfo()
    assert lib.hpcs_plan_query(None, N.HPCS_DTYPE_U32, 1 << 28, 0, ctypes.byref(info)) == 0
    assert info.passes == 4 and info.digit_bits == 8
    assert info.ideal_bytes == (1 << 28) * 36     # one histogram read + read and write per pass
    assert lib.hpcs_plan_query(None, 99, 10, 0, ctypes.byref(info)) == N.HPCS_E_INVALID_ARG


def _table(tmp_path, text):
    p = tmp_path / "tune.txt"
    p.write_text(text)
    return str(p).encode()


def test_tuning_table_parser(tmp_path):
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libhpcsort.so not built")
    lib = N.lib()
    rows = ctypes.c_int32()
    good = ("# comment\nu32 4096 0 0 1   # best\n\nf64 268435456 1 0 1\npairs64 1024 0 0 0\nu64 65536 0 0 0 11\n"
            "i32 1048576 4 4 1\npairs64 65536 5 0 1\nu32 268435456 6 0 3\nsegs 268435456 4096 2 9\n")
    assert lib.hpcs_tuning_check(_table(tmp_path, good), ctypes.byref(rows)) == 0
    assert rows.value == 8
    for bad in ("u16 4096 0 0 1\n", "u64 4096 0 2 1\n", "u32 4096 0 0 1 7\n", "u32 0 0 0 1\n",
                "u32 4096 2 0 1\n", "u32 4096 3 0 0 11\n", "pairs32 100 0 1 0\n", "u32 4096 0 0\n", "u32 4096 1 0 0 11\n",
                "pairs64 4096 0 0 0 11\n", "pairs32 4096 4 0 1\n", "u32 4096 7 0 1\n", "u64 4096 6 0 0\n",
                "pairs32 4096 6 0 1\n", "u32 4096 6 0 0 11\n", "u64 4096 4 0 0 11\n",
                "u32 4096 5 0 0 11\n", "segs 4096 4096 2 5\n"):
        assert lib.hpcs_tuning_check(_table(tmp_path, bad), ctypes.byref(rows)) == N.HPCS_E_INVALID_ARG, bad
        assert b"line 1" in lib.hpcs_last_error()
    assert lib.hpcs_tuning_check(b"/nonexistent/tune.txt", None) == N.HPCS_E_INVALID_ARG


def test_committed_tuning_table_parses():
    if not os.path.exists(N.TUNE_PATH):
        pytest.skip("no tuning table committed")
    rows = ctypes.c_int32()
    assert N.lib().hpcs_tuning_check(N.TUNE_PATH.encode(), ctypes.byref(rows)) == 0
    assert rows.value > 0


def test_nbit_sort_argument_checks():
    """SortRadixLsdNbit rejects digit widths beyond the key (RadixSortLsd.cs:
    1621-1622, 1844-1845) or below 1, and a null array, before any device
    work."""
    import numpy as np
    from hpcsharp_amd.algorithm import Algorithm, ArgumentNullException, ArgumentOutOfRangeException
    with pytest.raises(ArgumentNullException):
        Algorithm.SortRadixLsdNbit(None, 0, 0, 8)
    for arr, bad in ((np.zeros(4, np.uint32), 33), (np.zeros(4, np.uint32), 0), (np.zeros(4, np.uint64), 65)):
        with pytest.raises(ArgumentOutOfRangeException):
            Algorithm.SortRadixLsdNbit(arr, 0, 4, bad)
    with pytest.raises(ArgumentNullException):
        Algorithm.SortRadixLsd11bit(None)


def test_partition_radix_argument_checks():
    """PartitionRadix (PartitionRadix.cs:135-210): null array or k ->
    ArgumentNullException, empty array -> ArgumentOutOfRangeException, more
    indices than keys -> ArgumentException; all before any device work."""
    import numpy as np
    from hpcsharp_amd.algorithm import (Algorithm, ArgumentException, ArgumentNullException,
                                        ArgumentOutOfRangeException)
    with pytest.raises(ArgumentNullException):
        Algorithm.PartitionRadix(None, 3)
    with pytest.raises(ArgumentNullException):
        Algorithm.PartitionRadix(np.zeros(4, np.uint32), 0, 4, None)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.PartitionRadix(np.zeros(0, np.uint32), 0)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.PartitionRadix(np.zeros(4, np.uint32), 0, 0, [1])
    with pytest.raises(ArgumentException):
        Algorithm.PartitionRadix(np.zeros(2, np.uint32), [0, 1, 1])
    a = np.array([3, 1, 2], np.uint32)
    Algorithm.PartitionRadix(a, [])            # no indices: unchanged, no device call
    assert a.tolist() == [3, 1, 2]
    assert issubclass(ArgumentNullException, ArgumentException)
    assert issubclass(ArgumentOutOfRangeException, ArgumentException)


def test_tile_shapes_dispatch_agrees_with_the_tuning_parser(tmp_path):
    """Every tile shape the kernel dispatch accepts for a sort runs all of
    that sort's pass kernels (upsweep, tile counts, scatter / onesweep) with
    one (threads, keys per thread) tile; a shape the dispatch rejects is also
    rejected in a tuning row for that sort, and the other way round.  No
    device is used."""
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libhpcsort.so not built")
    lib = N.lib()
    kinds = {(4, 0): "u32", (4, 1): "pairs32", (8, 0): "u64", (8, 1): "pairs64"}
    served = {k: [] for k in kinds}
    for shape in range(-1, 12):
        for (kb, v), dtype in kinds.items():
            b, i = ctypes.c_int32(0), ctypes.c_int32(0)
            rc = lib.hpcs_tile_shape_query(shape, kb, v, ctypes.byref(b), ctypes.byref(i))
            row = tmp_path / f"row_{shape}_{dtype}.txt"
            row.write_text(f"{dtype} 1024 1 {shape} 0\n")
            rows = ctypes.c_int32(0)
            parsed = lib.hpcs_tuning_check(str(row).encode(), ctypes.byref(rows)) == 0
            if rc == N.HPCS_OK:
                served[(kb, v)].append(shape)
                assert b.value in (256, 512, 1024) and i.value in (8, 12, 16, 20, 24, 32), (shape, dtype)
                # the staging area of a tile fits one workgroup's LDS budget
                assert b.value * i.value * (kb + 4 * v) <= 64 * 1024, (shape, dtype)
                assert parsed, (shape, dtype)
            else:
                assert rc == N.HPCS_E_UNSUPPORTED, (shape, dtype, rc)
                assert not parsed, (shape, dtype)
    # shape 0 serves every sort; the 1024-thread rows exist
    assert all(0 in s for s in served.values())
    assert 5 in served[(4, 0)] and 6 in served[(8, 0)] and 6 in served[(4, 1)]
    b, i = ctypes.c_int32(0), ctypes.c_int32(0)
    assert lib.hpcs_tile_shape_query(0, 3, 0, ctypes.byref(b), ctypes.byref(i)) == N.HPCS_E_INVALID_ARG


# ---- the C# binding in INTEGRATION.md ----------------------------------------

_CS_WIDTH = {"int": ("int", 4), "uint": ("int", 4), "long": ("int", 8), "ulong": ("int", 8)}
_C_WIDTH = {"int32_t": ("int", 4), "int64_t": ("int", 8), "uint32_t": ("int", 4), "uint64_t": ("int", 8),
            "int": ("int", 4)}


def _cs_param_kind(p: str):
    p = p.strip()
    toks = p.replace("*", " * ").split()
    if toks[0] == "out" or "*" in toks or toks[0] == "IntPtr":
        return ("ptr", 8)
    return _CS_WIDTH[toks[0]]


def _c_param_kind(p: str):
    p = p.strip()
    if "*" in p:
        return ("ptr", 8)
    toks = [t for t in p.split() if t != "const"]
    return _C_WIDTH[toks[0]]


def _c_prototypes() -> dict[str, list]:
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:int32_t|const char\s*\*)\s*(hpcs_\w+)\s*\(([^)]*)\)\s*;", text):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        protos[m.group(1)] = [_c_param_kind(a) for a in args]
    return protos


def _csharp_source() -> str:
    with open(os.path.join(REPO, "INTEGRATION.md")) as f:
        text = f.read()
    blocks = re.findall(r"```csharp\n(.*?)```", text, flags=re.S)
    assert blocks, "no C# block in INTEGRATION.md"
    return "\n".join(blocks)


def test_csharp_library_imports_match_the_header():
    """Every [LibraryImport] prototype of the C# binding has the header's
    parameter count, and at each position a pointer for a pointer and an
    integer of the same width for an integer."""
    protos = _c_prototypes()
    src = _csharp_source()
    imports = re.findall(r"\[LibraryImport\(Lib\)\]\s*internal static (?:unsafe )?partial (\w+) (hpcs_\w+)\(([^)]*)\);",
                         src)
    assert len(imports) >= 20, len(imports)
    for ret, name, params in imports:
        assert name in protos, name
        cs = [_cs_param_kind(p) for p in params.split(",") if p.strip()]
        assert cs == protos[name], (name, cs, protos[name])
        assert ret in ("int", "IntPtr"), (name, ret)


def _normalize_cs_type(t: str) -> str:
    t = re.sub(r"\s+", "", t)
    for a, b in (("UInt32", "uint"), ("UInt64", "ulong"), ("Int32", "int"), ("Int64", "long")):
        t = re.sub(rf"\b{a}\b", b, t)
    return t


def _cs_methods(src: str) -> set[tuple]:
    out = set()
    for m in re.finditer(r"public static [^(=]*?\b(\w+)(?:<T>)?\(([^)]*)\)", src):
        params = []
        depth, cur = 0, ""
        for ch in m.group(2):
            if ch == "<":
                depth += 1
            elif ch == ">":
                depth -= 1
            if ch == "," and depth == 0:
                params.append(cur)
                cur = ""
            else:
                cur += ch
        if cur.strip():
            params.append(cur)
        types = []
        for p in params:
            p = p.split("=")[0].strip()
            if p.startswith("this "):
                p = p[5:]
            types.append(_normalize_cs_type(p.rsplit(" ", 1)[0]))
        out.add((m.group(1), tuple(types)))
    return out


# The reference's drop-in surface for this path (HPCsharp/RadixSortLsd.cs,
# RadixSortLsdParallel.cs, Histogram.cs, HistogramParallel.cs; line of each):
# name and parameter types, which a caller's code binds to.
REFERENCE_OVERLOADS = [
    ("SortRadixLsd", ("uint[]",)),                                   # RadixSortLsd.cs:255
    ("SortRadixLsd", ("uint[]", "int", "int")),                      # :242
    ("SortRadixLsdFunc", ("uint[]",)),                               # :227
    ("SortRadixLsdFunc", ("uint[]", "int", "int")),                  # :213
    ("SortRadixLsd", ("int[]",)), ("SortRadixLsd", ("int[]", "int", "int")),          # :352, :339
    ("SortRadixLsdFunc", ("int[]",)), ("SortRadixLsdFunc", ("int[]", "int", "int")),  # :324, :310
    ("SortRadixLsd", ("ulong[]",)), ("SortRadixLsd", ("ulong[]", "int", "int")),      # :437, :424
    ("SortRadixLsdFunc", ("ulong[]",)), ("SortRadixLsdFunc", ("ulong[]", "int", "int")),  # :409, :395
    ("SortRadixLsd", ("long[]",)), ("SortRadixLsd", ("long[]", "int", "int")),        # :527, :514
    ("SortRadixLsdFunc", ("long[]",)), ("SortRadixLsdFunc", ("long[]", "int", "int")),    # :500, :486
    ("SortRadixLsd", ("List<uint>",)),                               # :615
    ("SortRadixLsd", ("T[]", "Func<T,uint>")),                       # :714
    ("SortRadixLsd2", ("T[]", "Func<T,uint>")),                      # :751
    ("SortRadixLsd", ("T[]", "Func<T,ulong>")),                      # :783
    ("SortRadixLsd2", ("T[]", "Func<T,ulong>")),                     # :821
    ("SortRadixFaster", ("T[]", "Func<T,uint>")),                    # :861
    ("SortRadixFaster2", ("T[]", "Func<T,uint>")),                   # :917
    ("SortRadixFaster", ("T[]", "Func<T,ulong>")),                   # :961
    ("SortRadixFaster2", ("T[]", "Func<T,ulong>")),                  # :1016
    ("SortRadixFaster", ("List<T>", "Func<T,uint>")),                # :1059
    ("SortRadixFaster", ("List<T>", "Func<T,ulong>")),               # :1075
    ("SortRadix", ("T[]", "int", "int", "T[]", "Func<T,uint>")),     # :1091
    ("SortRadix", ("List<T>", "Func<T,uint>")),                      # :1151
    ("SortRadix", ("List<T>", "Func<T,ulong>")),                     # :1167
    ("SortRadix", ("T[]", "uint[]")),                                # :1183
    ("SortRadix", ("T[]", "ulong[]")),                               # :1245
    ("SortRadix", ("List<T>", "uint[]")),                            # :1307
    ("SortRadix", ("List<T>", "ulong[]")),                           # :1323
    ("SortRadixLsd11bit", ("uint[]",)),                              # :1540
    ("SortRadixLsdNbit", ("uint[]", "int", "int", "int")),           # :1616
    ("SortRadixLsdNbit", ("ulong[]", "int", "int", "int")),          # :1839
    ("SortRadixPar", ("uint[]", "int")),                             # RadixSortLsdParallel.cs:145
    ("SortRadixPar", ("int[]",)),                                    # :446
    ("SortRadixPar", ("long[]",)),                                   # :560
    ("SortRadixPar", ("ulong[]",)),                                  # :690
    ("HistogramByteComponents", ("uint[]", "int", "int")),           # Histogram.cs:261
    ("Histogram11bitComponents", ("uint[]", "int", "int")),          # :298
    ("HistogramNbitComponents", ("uint[]", "int", "int", "int")),    # :381
    ("HistogramNbitComponents", ("ulong[]", "int", "int", "int")),   # :405
    ("HistogramByteComponents", ("T[]", "int", "int", "Func<T,uint>")),    # :760
    ("HistogramByteComponents", ("T[]", "int", "int", "Func<T,ulong>")),   # :799
    ("HistogramByteComponentsAndKeyArray", ("T[]", "int", "int", "Func<T,uint>")),   # :846
    ("HistogramByteComponentsAndKeyArray", ("T[]", "int", "int", "Func<T,ulong>")),  # :887
    ("HistogramByteComponents", ("ulong[]", "int", "int")),          # :957
    ("HistogramByteComponents", ("int[]", "int", "int")),            # :992
    ("HistogramByteComponents", ("long[]", "int", "int")),           # :1019
    ("HistogramByteComponentsPar", ("uint[]", "int", "int", "int")),   # HistogramParallel.cs:214
    ("HistogramByteComponentsPar", ("int[]", "int", "int", "int")),    # :416
    ("HistogramByteComponentsPar", ("ulong[]", "int", "int", "int")),  # :547
    ("HistogramByteComponentsPar", ("long[]", "int", "int", "int")),   # :617
    ("SelectTopK", ("uint[]", "int", "int", "int")),                 # SelectionTopKRadix.cs:142
    ("SelectTopK", ("uint[]", "int")),                               # :160
    ("SortRadixMsd", ("float[]",)), ("SortRadixMsd", ("double[]",)),   # RadixSortMsd.cs:1563, :1587
]


def test_csharp_binding_covers_the_reference_overloads():
    """Every overload of the reference's surface for this path has a binding
    method of the same name and parameter types, so a caller's code (e.g.
    `a.SortRadixLsd2(e => e.Key)`) compiles against HPCsharp.Gpu unchanged."""
    have = _cs_methods(_csharp_source())
    missing = [m for m in REFERENCE_OVERLOADS if m not in have]
    assert not missing, missing
