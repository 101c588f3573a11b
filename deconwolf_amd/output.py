"""Output path of dw_run: the statistics, 16-bit scaling and file writers
deconwolf applies to a method's result (src/dw.c:2862-2904 for a whole
image, src/dw.c:2162-2200 after tiles).

The per-voxel work (largest value, voxels at the background level, the
float -> uint16 conversion of each writer) runs on the GPU through
libdwhip (dwhip_output_stats_device, dwhip_to_u16_device); the files are
written here:

* .npy: float32 or uint16, C order, shape (P, N, M) (npio_write,
  src/npio.c, through fim_imwrite_f32/_u16, src/fim.c:5108-5155);
* .tif: one uncompressed page per z plane, width M, height N, 16-bit
  unsigned or 32-bit float samples, min-is-black, BigTIFF when the 16-bit
  data reach 4 GiB (fim_tiff_write_opt, src/fim_tiff.c:1066-1170, and
  the float writer), optionally with deconwolf's ImageJ description,
  resolution and software tags (ttags_set, src/fim_tiff.c:1483-1513).

Rounding of the uint16 writers (all after multiplying by the scaling in
float): TIFF of a whole image rounds half away from zero after clamping
to 65535 (non-finite -> 0); .npy of a whole image and TIFF after tiles
truncate; .npy after tiles rounds half to even (nearbyintf).
"""
import ctypes
import struct

import numpy as np

from ._lib import U16_RINT, U16_ROUND, U16_TRUNC, check, lib  # noqa: F401  (re-exported)


def scaling_for_u16(vmax):
    """scaling_for_u16 (src/dw.c:98-103): (2^16 - 1) / max in double,
    returned as float."""
    return float(np.float32((2.0 ** 16 - 1.0) / float(np.float32(vmax))))


def tiled_scaling(vmax, scaling):
    """The scaling deconvolve_tiles picks (src/dw.c:2162-2172): kept when
    positive, else 65535 / max (float) when max > 0."""
    if scaling <= 0 and vmax > 0:
        return float(np.float32(65535) / np.float32(vmax))
    return float(scaling)


def stats(out, bg):
    """(largest value, voxels equal to bg) of a float32 GPU tensor."""
    import torch
    out = out.contiguous()
    vmax = ctypes.c_float()
    nbg = ctypes.c_int64()
    check(lib().dwhip_output_stats_device(out.data_ptr(), out.numel(), float(bg), ctypes.byref(vmax),
                                          ctypes.byref(nbg), torch.cuda.current_stream().cuda_stream),
          "dwhip_output_stats_device")
    return float(vmax.value), int(nbg.value)


def to_u16(out, scaling, mode):
    """uint16 volume (numpy, same shape) that deconwolf's writer `mode`
    (U16_ROUND, U16_TRUNC, U16_RINT) stores for the float32 GPU tensor."""
    import torch
    out = out.contiguous()
    dst = torch.empty(out.shape, dtype=torch.int16, device=out.device)
    check(lib().dwhip_to_u16_device(out.data_ptr(), out.numel(), float(scaling), int(mode), dst.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream), "dwhip_to_u16_device")
    return dst.cpu().numpy().view(np.uint16)


# ---------------------------------------------------------------------------
# file writers (host)
# ---------------------------------------------------------------------------
def write_npy(path, vol):
    np.save(path, np.ascontiguousarray(vol), allow_pickle=False)


def imagej_description(P, dz_nm):
    """ttags_set_pixelsize's ImageJ description (src/fim_tiff.c:1279-1281)."""
    return "ImageJ=1.52r\nimages=%d\nslices=%d\nunit=nm\nspacing=%.1f\nloop=false" % (P, P, dz_nm)


_SHORT, _LONG, _RATIONAL, _ASCII, _LONG8 = 3, 4, 5, 2, 16


def _rational(x):
    # the largest denominator that keeps the numerator in 32 bits
    den = 1
    while den < (1 << 20) and abs(x) * den * 2 < (1 << 32) - 1:
        den *= 2
    return int(round(x * den)), den


def write_tiff(path, vol, description=None, software=None, resolution=None, bigtiff=None):
    """Multi-page TIFF of a (P, N, M) uint16 or float32 volume.

    resolution: pixels per unit in x and y (ttags' xresolution), written
    with ResolutionUnit 1 (none), as deconwolf's tags carry it."""
    vol = np.ascontiguousarray(vol)
    if vol.ndim != 3:
        raise ValueError("vol must be (P, N, M)")
    if vol.dtype == np.uint16:
        bps, sfmt = 16, 1
    elif vol.dtype == np.float32:
        bps, sfmt = 32, 3
    else:
        raise ValueError("uint16 or float32 volumes only")
    P, N, M = vol.shape
    if bigtiff is None:
        bigtiff = vol.nbytes >= (1 << 32)
    # layout: header, then per page the strip followed by its IFD and the
    # IFD's out-of-line values
    ofmt, osz = ("<Q", 8) if bigtiff else ("<I", 4)
    cfmt = "<Q" if bigtiff else "<H"
    esz = 20 if bigtiff else 12
    plane = M * N * vol.itemsize
    with open(path, "wb") as f:
        if bigtiff:
            f.write(b"II" + struct.pack("<HHHQ", 43, 8, 0, 0))
            first_ptr = 8
        else:
            f.write(b"II" + struct.pack("<HI", 42, 0))
            first_ptr = 4
        prev_ptr = first_ptr
        for z in range(P):
            data_off = f.tell()
            f.write(vol[z].tobytes())
            entries = [
                (256, _LONG, 1, M), (257, _LONG, 1, N), (258, _SHORT, 1, bps), (259, _SHORT, 1, 1),
                (262, _SHORT, 1, 1),
            ]
            if description is not None:
                entries.append((270, _ASCII, None, description.encode("ascii") + b"\0"))
            entries += [
                (273, _LONG8 if bigtiff else _LONG, 1, data_off), (274, _SHORT, 1, 1), (277, _SHORT, 1, 1),
                (278, _LONG, 1, N), (279, _LONG8 if bigtiff else _LONG, 1, plane),
            ]
            if resolution is not None:
                r = _rational(float(resolution))
                entries += [(282, _RATIONAL, 1, r), (283, _RATIONAL, 1, r)]
            entries.append((284, _SHORT, 1, 1))
            if resolution is not None:
                entries.append((296, _SHORT, 1, 1))
            if software is not None:
                entries.append((305, _ASCII, None, software.encode("ascii") + b"\0"))
            entries.append((339, _SHORT, 1, sfmt))
            ifd_off = f.tell()
            if ifd_off % 2:
                f.write(b"\0")
                ifd_off += 1
            # patch the pointer to this IFD
            f.seek(prev_ptr)
            f.write(struct.pack(ofmt, ifd_off))
            f.seek(ifd_off)
            n = len(entries)
            head = struct.calcsize(cfmt)
            after = ifd_off + head + n * esz + osz          # out-of-line values go here
            body = bytearray()
            blob = bytearray()
            for tag, typ, cnt, val in entries:
                if typ == _ASCII:
                    raw, cnt = val, len(val)
                elif typ == _RATIONAL:
                    raw = struct.pack("<II", *val)
                elif typ == _SHORT:
                    raw = struct.pack("<H", val)
                elif typ == _LONG:
                    raw = struct.pack("<I", val)
                else:
                    raw = struct.pack("<Q", val)
                if len(raw) <= osz:
                    field = raw + b"\0" * (osz - len(raw))
                else:
                    if (after + len(blob)) % 2:
                        blob += b"\0"
                    field = struct.pack(ofmt, after + len(blob))
                    blob += raw
                body += struct.pack("<HH", tag, typ) + struct.pack(ofmt, cnt) + field
            f.write(struct.pack(cfmt, n) + bytes(body))
            prev_ptr = f.tell()
            f.write(struct.pack(ofmt, 0) + bytes(blob))
    return path


def _is_npy(path):
    return str(path).lower().endswith(".npy")


def imwrite(path, out, s, tiled=False, log=None):
    """Write a method result like dw_run: `out` is the (P, N, M) float32
    GPU tensor the method (or the tile blend) returned, `s` a DwOpts with
    outFormat (16 or 32), scaling (<= 0: automatic) and bg.  Returns the
    scaling used for 16-bit output (None for 32-bit).  Log lines as
    deconwolf writes them."""
    P, N, M = out.shape
    vmax, nbg = stats(out, s.bg)
    if not tiled and log is not None:
        # dw.c:2869-2870
        log.write("%f%% pixels at bg level in the output image.\n" % (100.0 * nbg / (M * N * P)))
    if int(s.outFormat) == 32:
        if tiled:
            s.scaling = 1.0
            if log is not None:
                log.write("scaling: %f\n" % s.scaling)
        vol = out.contiguous().cpu().numpy()
        write_npy(path, vol) if _is_npy(path) else write_tiff(path, vol)
        return None
    if tiled:
        s.scaling = tiled_scaling(vmax, s.scaling)
        mode = U16_RINT if _is_npy(path) else U16_TRUNC
    else:
        if s.scaling <= 0:
            s.scaling = scaling_for_u16(vmax)
        mode = U16_TRUNC if _is_npy(path) else U16_ROUND
    if log is not None:
        log.write("scaling: %f\n" % s.scaling)
    vol = to_u16(out, s.scaling, mode)
    write_npy(path, vol) if _is_npy(path) else write_tiff(path, vol)
    return s.scaling
