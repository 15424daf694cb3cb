"""The built library's GPU code, inspected on the CPU: every kernel of
libcapytaine_amd.so runs without scratch memory (private segment size 0 in
its kernel descriptor), so no dispatch depends on the runtime finding
scratch space on a nearly full device, and every entry point of
include/capytaine_amd.h is exported.

The code objects are read from the library's offload bundles (one per
translation unit): the gfx950 ELF of each, its symbol table, and for each
kernel descriptor `<kernel>.kd` (64 bytes, AMDHSA code object v5) the
private segment fixed size at byte 4."""
import os
import re
import struct

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "capytaine_amd", "libcapytaine_amd.so")
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _sections(elf):
    shoff = struct.unpack_from("<Q", elf, 0x28)[0]
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", elf, 0x3A)
    secs = []
    for i in range(shnum):
        name, typ, flags, addr, off, size, link, info, align, entsize = struct.unpack_from(
            "<IIQQQQIIQQ", elf, shoff + i * shentsize)
        secs.append(dict(name=name, type=typ, addr=addr, off=off, size=size, link=link, entsize=entsize))
    strtab = secs[shstrndx]
    for s in secs:
        raw = elf[strtab["off"] + s["name"]:]
        s["name"] = raw[:raw.index(b"\0")].decode()
    return secs


def _gfx950_objects(lib):
    """The gfx950 code objects of every offload bundle in .hip_fatbin."""
    sec = next(s for s in _sections(lib) if s["name"] == ".hip_fatbin")
    fat = lib[sec["off"]:sec["off"] + sec["size"]]
    out = []
    pos = fat.find(MAGIC)
    while pos >= 0:
        n = struct.unpack_from("<Q", fat, pos + 24)[0]
        p = pos + 32
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", fat, p)
            triple = fat[p + 24:p + 24 + tlen].decode()
            p += 24 + tlen
            if "gfx950" in triple:
                out.append(fat[pos + off:pos + off + size])
        pos = fat.find(MAGIC, pos + 1)
    return out


def _kernel_descriptors(obj):
    """{kernel: (group segment bytes, private segment bytes)}."""
    secs = _sections(obj)
    symtab = next(s for s in secs if s["type"] == 2)  # SHT_SYMTAB
    strtab = secs[symtab["link"]]
    out = {}
    for i in range(symtab["size"] // symtab["entsize"]):
        name_off, info, other, shndx, value, size = struct.unpack_from(
            "<IBBHQQ", obj, symtab["off"] + i * symtab["entsize"])
        raw = obj[strtab["off"] + name_off:]
        name = raw[:raw.index(b"\0")].decode()
        if not name.endswith(".kd") or shndx >= len(secs):
            continue
        s = secs[shndx]
        fo = s["off"] + (value - s["addr"])
        group, private = struct.unpack_from("<II", obj, fo)
        out[name[:-3]] = (group, private)
    return out


@pytest.fixture(scope="module")
def lib_bytes():
    if not os.path.exists(LIB):
        pytest.skip("libcapytaine_amd.so is not built")
    return open(LIB, "rb").read()


def test_every_kernel_runs_without_scratch(lib_bytes):
    objs = _gfx950_objects(lib_bytes)
    assert len(objs) == 7  # one per translation unit
    kernels = {}
    for o in objs:
        kernels.update(_kernel_descriptors(o))
    assert len(kernels) > 50
    assert any("panel_mw_kernel" in k for k in kernels) and any("panel_kernel" in k for k in kernels)
    scratch = {k: v[1] for k, v in kernels.items() if v[1] != 0}
    assert not scratch, f"kernels with scratch (bytes per lane): {scratch}"


def test_every_declared_entry_point_is_exported(lib_bytes):
    from capytaine_amd import _lib
    text = open(os.path.join(ROOT, "include", "capytaine_amd.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(cpt_\w+)\s*\(", text, flags=re.M))
    assert "cpt_lu_solve_noalloc" in declared and "cpt_mixed_refine" in declared
    lib = _lib.load()
    for name in declared:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name
