"""Register, scratch and LDS figures of the library's own gfx950 kernels, read
from the AMDGPU metadata note of the device code object inside each built
translation unit (finitevolume_amd/_build/*.o: the clang offload bundle in
its .hip_fatbin section).  Test infrastructure: it reads only objects this
repository builds from its own sources."""
import os
import struct

import msgpack

BUILD = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "finitevolume_amd", "_build")
_BUNDLE_MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
_NT_AMDGPU_METADATA = 32


def _sections(elf: bytes):
    """(name, type, offset, size) of each section of a little-endian ELF64 image."""
    if elf[:4] != b"\x7fELF" or elf[4] != 2 or elf[5] != 1:
        raise ValueError("not a little-endian ELF64 image")
    shoff = struct.unpack_from("<Q", elf, 0x28)[0]
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", elf, 0x3A)
    hdrs = []
    for i in range(shnum):
        name, typ, _flags, _addr, off, size = struct.unpack_from("<IIQQQQ", elf, shoff + i * shentsize)
        hdrs.append((name, typ, off, size))
    stroff = hdrs[shstrndx][2]
    out = []
    for name, typ, off, size in hdrs:
        end = elf.index(b"\0", stroff + name)
        out.append((elf[stroff + name:end].decode(), typ, off, size))
    return out


def _device_image(obj: bytes, arch: str = "gfx950") -> bytes:
    for name, _typ, off, size in _sections(obj):
        if name == ".hip_fatbin":
            fat = obj[off:off + size]
            break
    else:
        raise ValueError("no .hip_fatbin section")
    if fat[:24] != _BUNDLE_MAGIC:
        raise ValueError("unsupported offload bundle format")
    n = struct.unpack_from("<Q", fat, 24)[0]
    p = 32
    for _ in range(n):
        eoff, esize, tlen = struct.unpack_from("<QQQ", fat, p)
        p += 24
        triple = fat[p:p + tlen].decode()
        p += tlen
        if triple.endswith(arch):
            return fat[eoff:eoff + esize]
    raise ValueError(f"no {arch} code object in the bundle")


def _metadata(image: bytes) -> dict:
    for _name, typ, off, size in _sections(image):
        if typ != 7:  # SHT_NOTE
            continue
        p, end = off, off + size
        while p + 12 <= end:
            namesz, descsz, ntype = struct.unpack_from("<III", image, p)
            p += 12
            name = image[p:p + namesz].rstrip(b"\0")
            p += (namesz + 3) & ~3
            desc = image[p:p + descsz]
            p += (descsz + 3) & ~3
            if name == b"AMDGPU" and ntype == _NT_AMDGPU_METADATA:
                return msgpack.unpackb(desc, raw=False)
    raise ValueError("no AMDGPU metadata note")


def kernel_resources(unit: str) -> dict:
    """{kernel symbol: {"vgpr", "sgpr", "scratch", "lds"}} of a built
    translation unit (e.g. "kernels_euler")."""
    with open(os.path.join(BUILD, unit + ".o"), "rb") as f:
        meta = _metadata(_device_image(f.read()))
    out = {}
    for k in meta["amdhsa.kernels"]:
        out[k[".name"]] = {"vgpr": k[".vgpr_count"], "sgpr": k[".sgpr_count"],
                           "scratch": k[".private_segment_fixed_size"], "lds": k[".group_segment_fixed_size"]}
    return out
