"""CPU ORACLE — TEST INFRASTRUCTURE ONLY (JPEG video frames).

Restates the JPEG path of serverSendFrame (lib/isaac/isaac_communicator.hpp:221-250):
libjpeg-turbo compressing the RGBX frame with JCS_EXT_RGBX input, jpeg_set_defaults
and jpeg_set_quality(quality, FALSE), then the base64 payload
{"payload": "data:image/jpeg;base64,..."}.  libjpeg-turbo is not in the reference
(a system dependency, found by lib/ISAACConfig.cmake:81-101); its published
baseline algorithm as jpeg_set_defaults configures it is restated here:

  jcparam.c   quality scaling (q < 50: 5000 / q, else 200 - 2q; entries
              (t * scale + 50) / 100 clamped to [1, 32767]) of the T.81 Annex K
              tables; Y sampled 2x2, Cb and Cr 1x1 (4:2:0); standard Huffman tables
  jccolor.c   RGB -> YCbCr in 16-bit fixed point
  jcsample.c  fullsize copy (Y) and 2x2 box with alternating bias 1, 2 (Cb, Cr),
              the right edge replicated to whole blocks
  jcprepct.c  the bottom edge replicated to whole MCU rows
  jfdctint.c  the accurate integer DCT (LL&M, 13-bit constants, 2 extra bits in pass 1)
  jcdctmgr.c  quantisation of DCT/8 by rounding division
  jccoefct.c  dummy blocks past the image (zero AC, DC of the block before)
  jchuff.c    DC differences and AC run lengths coded with the tables, 0xFF
              stuffing, 1-bit padding
  jcmarker.c  SOI, JFIF APP0, DQT per table, SOF0, DHT per table, SOS, EOI

The tables come from tests/golden/jpeg_tables.json (T.81 Annex K).  Pillow's
bundled libjpeg-turbo pins this restatement byte for byte (tests/test_jpeg_oracle.py).
Only tests/ may import this module; the product never does.
"""
from __future__ import annotations

import base64
import json
import os
from typing import Dict, List, Tuple

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden", "jpeg_tables.json")


def tables():
    return json.load(open(GOLDEN))


def quality_tables(quality: int, force_baseline: bool = False) -> Tuple[np.ndarray, np.ndarray]:
    """natural-order luminance and chrominance tables for jpeg_set_quality(q)"""
    t = tables()
    q = min(max(int(quality), 1), 100)
    scale = 5000 // q if q < 50 else 200 - 2 * q
    out = []
    for k in ("0", "1"):
        base = np.asarray(t["quant_natural"][k], np.int64)
        v = np.clip((base * scale + 50) // 100, 1, 32767)
        if force_baseline:
            v = np.minimum(v, 255)
        out.append(v)
    return out[0], out[1]


def _fix(x: float) -> int:
    return int(x * 65536 + 0.5)


def rgb_to_ycc(rgb: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    r, g, b = (rgb[..., i].astype(np.int64) for i in range(3))
    half, off = 1 << 15, 128 << 16
    y = (_fix(0.29900) * r + _fix(0.58700) * g + _fix(0.11400) * b + half) >> 16
    cb = (-_fix(0.16874) * r - _fix(0.33126) * g + _fix(0.5) * b + off + half - 1) >> 16
    cr = (_fix(0.5) * r - _fix(0.41869) * g - _fix(0.08131) * b + off + half - 1) >> 16
    return y, cb, cr


def _expand(p: np.ndarray, rows: int, cols: int) -> np.ndarray:
    """replicate the last column / row out to (rows, cols)"""
    h, w = p.shape
    if cols > w:
        p = np.concatenate([p, np.repeat(p[:, -1:], cols - w, axis=1)], axis=1)
    if rows > h:
        p = np.concatenate([p, np.repeat(p[-1:, :], rows - h, axis=0)], axis=0)
    return p


def planes(rgb: np.ndarray):
    """the three component planes as the coefficient controller sees them, and
    each component's size in blocks (width_in_blocks, height_in_blocks)"""
    H, W = rgb.shape[:2]
    y, cb, cr = rgb_to_ycc(rgb)
    mcux, mcuy = -(-W // 16), -(-H // 16)
    # Y: fullsize copy, right edge to whole blocks; bottom to whole MCU rows
    ybw, ybh = -(-W // 8), -(-H // 8)
    # the conversion buffer holds 2 rows: an odd last row group is completed by its last row
    Hg = H + (H % 2)
    yp = _expand(y, Hg, ybw * 8)
    yp = _expand(yp, mcuy * 16, ybw * 8)
    # chroma: 2x2 box with bias 1, 2 over the right-expanded full-size plane
    cw, ch = -(-W // 2), -(-H // 2)
    cbw, cbh = -(-cw // 8), -(-ch // 8)
    out = []
    for c in (cb, cr):
        p = _expand(c, Hg, cbw * 16)
        s = p[0:Hg:2, 0::2] + p[0:Hg:2, 1::2] + p[1:Hg:2, 0::2] + p[1:Hg:2, 1::2]
        bias = np.tile(np.array([1, 2], np.int64), s.shape[1] // 2 + 1)[: s.shape[1]]
        d = (s + bias[None, :]) >> 2
        out.append(_expand(d, mcuy * 8, cbw * 8))
    return [yp, out[0], out[1]], [(ybw, ybh), (cbw, cbh), (cbw, cbh)], (mcux, mcuy)


def fdct_islow(blocks: np.ndarray) -> np.ndarray:
    """(..., 8, 8) level-shifted samples -> the integer DCT scaled by 8 (jfdctint.c)"""
    CB, P1 = 13, 2

    def c(x):
        return int(x * (1 << CB) + 0.5)

    f0298, f0390, f0541, f0765 = c(0.298631336), c(0.390180644), c(0.541196100), c(0.765366865)
    f0899, f1175, f1501, f1847 = c(0.899976223), c(1.175875602), c(1.501321110), c(1.847759065)
    f1961, f2053, f2562, f3072 = c(1.961570560), c(2.053119869), c(2.562915447), c(3.072711026)

    def descale(x, n):
        return (x + (1 << (n - 1))) >> n

    def one_pass(d, first):
        d = [d[..., k] for k in range(8)]
        t0, t7 = d[0] + d[7], d[0] - d[7]
        t1, t6 = d[1] + d[6], d[1] - d[6]
        t2, t5 = d[2] + d[5], d[2] - d[5]
        t3, t4 = d[3] + d[4], d[3] - d[4]
        t10, t13 = t0 + t3, t0 - t3
        t11, t12 = t1 + t2, t1 - t2
        o = [None] * 8
        if first:
            o[0] = (t10 + t11) << P1
            o[4] = (t10 - t11) << P1
            sh = CB - P1
        else:
            o[0] = descale(t10 + t11, P1)
            o[4] = descale(t10 - t11, P1)
            sh = CB + P1
        z1 = (t12 + t13) * f0541
        o[2] = descale(z1 + t13 * f0765, sh)
        o[6] = descale(z1 + t12 * (-f1847), sh)
        z1, z2, z3, z4 = t4 + t7, t5 + t6, t4 + t6, t5 + t7
        z5 = (z3 + z4) * f1175
        t4, t5, t6, t7 = t4 * f0298, t5 * f2053, t6 * f3072, t7 * f1501
        z1, z2, z3, z4 = z1 * (-f0899), z2 * (-f2562), z3 * (-f1961), z4 * (-f0390)
        z3 = z3 + z5
        z4 = z4 + z5
        o[7] = descale(t4 + z1 + z3, sh)
        o[5] = descale(t5 + z2 + z4, sh)
        o[3] = descale(t6 + z2 + z3, sh)
        o[1] = descale(t7 + z1 + z4, sh)
        return np.stack(o, axis=-1)

    x = blocks.astype(np.int64)
    x = one_pass(x, True)  # rows
    x = np.swapaxes(one_pass(np.swapaxes(x, -1, -2), False), -1, -2)  # columns
    return x


def quantize(coef: np.ndarray, qnat: np.ndarray) -> np.ndarray:
    q = (qnat.astype(np.int64) << 3).reshape(8, 8)
    a = np.abs(coef)
    v = (a + (q >> 1)) // q
    return np.where(coef < 0, -v, v)


def _huff_codes(bits: List[int], vals: List[int]) -> Dict[int, Tuple[int, int]]:
    codes, code, k = {}, 0, 0
    for length in range(1, 17):
        for _ in range(bits[length - 1]):
            codes[vals[k]] = (code, length)
            code += 1
            k += 1
        code <<= 1
    return codes


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, code: int, length: int):
        self.acc = (self.acc << length) | (code & ((1 << length) - 1))
        self.n += length
        while self.n >= 8:
            self.n -= 8
            b = (self.acc >> self.n) & 0xFF
            self.out.append(b)
            if b == 0xFF:
                self.out.append(0)
        self.acc &= (1 << self.n) - 1

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)
        return bytes(self.out)


def block_coefficients(rgb: np.ndarray, quality: int):
    """quantised zigzag coefficients of every block in scan order (MCU by MCU:
    Y00 Y01 Y10 Y11 Cb Cr), dummy blocks included: (n_blocks, 64) and the
    component of each block."""
    t = tables()
    zz = np.asarray(t["zigzag"])
    ql, qc = quality_tables(quality)
    ps, sizes, (mx, my) = planes(rgb)
    comp = []
    for ci, p in enumerate(ps):
        h, w = p.shape
        b = (p - 128).reshape(h // 8, 8, w // 8, 8).swapaxes(1, 2)  # (by, bx, 8, 8)
        qv = quantize(fdct_islow(b), ql if ci == 0 else qc)
        comp.append(qv.reshape(qv.shape[:2] + (64,))[..., zz])  # zigzag
    out, cid = [], []
    for my_ in range(my):
        for mx_ in range(mx):
            for ci in range(3):
                bw, bh = sizes[ci]
                n = 2 if ci == 0 else 1
                last = None
                for v in range(n):
                    for u in range(n):
                        bx, by = mx_ * n + u, my_ * n + v
                        if bx < bw and by < bh:
                            blk = comp[ci][by, bx].copy()
                        else:  # dummy block: zero AC, DC of the block before it in the MCU
                            blk = np.zeros(64, np.int64)
                            blk[0] = last[0]
                        out.append(blk)
                        cid.append(ci)
                        last = blk
    return np.asarray(out), np.asarray(cid)


def entropy_code(blocks: np.ndarray, cid: np.ndarray) -> bytes:
    t = tables()
    h = t["huffman"]
    dc = [_huff_codes(h["dc0"]["bits"], h["dc0"]["vals"]), _huff_codes(h["dc1"]["bits"], h["dc1"]["vals"])]
    ac = [_huff_codes(h["ac0"]["bits"], h["ac0"]["vals"]), _huff_codes(h["ac1"]["bits"], h["ac1"]["vals"])]
    bw = _Bits()
    pred = [0, 0, 0]
    for blk, ci in zip(blocks.tolist(), cid.tolist()):
        tbl = 0 if ci == 0 else 1
        diff = blk[0] - pred[ci]
        pred[ci] = blk[0]
        nb = int(abs(diff)).bit_length()
        bw.put(*dc[tbl][nb])
        if nb:
            bw.put(diff if diff >= 0 else diff - 1, nb)
        run = 0
        for k in range(1, 64):
            v = blk[k]
            if v == 0:
                run += 1
                continue
            while run > 15:
                bw.put(*ac[tbl][0xF0])
                run -= 16
            nb = int(abs(v)).bit_length()
            bw.put(*ac[tbl][(run << 4) | nb])
            bw.put(v if v >= 0 else v - 1, nb)
            run = 0
        if run:
            bw.put(*ac[tbl][0x00])
    return bw.flush()


def headers(width: int, height: int, quality: int) -> bytes:
    t = tables()
    zz = t["zigzag"]
    ql, qc = quality_tables(quality)
    seg = bytearray(b"\xff\xd8")
    seg += b"\xff\xe0\x00\x10JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00"
    wide = False
    for k, q in ((0, ql), (1, qc)):
        if int(q.max()) > 255:  # 16-bit entries (jpeg_set_quality with force_baseline FALSE)
            wide = True
            seg += b"\xff\xdb\x00\x83" + bytes([0x10 | k]) + b"".join(int(q[zz[j]]).to_bytes(2, "big") for j in range(64))
        else:
            seg += b"\xff\xdb\x00\x43" + bytes([k]) + bytes(int(q[zz[j]]) for j in range(64))
    # SOF0 (baseline) unless a table needs 16 bits: SOF1 (extended sequential)
    seg += (b"\xff\xc1" if wide else b"\xff\xc0") + b"\x00\x11\x08" + height.to_bytes(2, "big") + width.to_bytes(2, "big")
    seg += b"\x03\x01\x22\x00\x02\x11\x01\x03\x11\x01"
    for name, cls in (("dc0", 0x00), ("ac0", 0x10), ("dc1", 0x01), ("ac1", 0x11)):
        bits, vals = t["huffman"][name]["bits"], t["huffman"][name]["vals"]
        seg += b"\xff\xc4" + (3 + 16 + len(vals)).to_bytes(2, "big") + bytes([cls]) + bytes(bits) + bytes(vals)
    seg += b"\xff\xda\x00\x0c\x03\x01\x00\x02\x11\x03\x11\x00\x3f\x00"
    return bytes(seg)


def encode(rgb: np.ndarray, quality: int = 90) -> bytes:
    """(H, W, 3 or 4) uint8 -> the JPEG file libjpeg-turbo writes (the 4th channel ignored)"""
    rgb = np.ascontiguousarray(rgb[..., :3])
    blocks, cid = block_coefficients(rgb, quality)
    return headers(rgb.shape[1], rgb.shape[0], quality) + entropy_code(blocks, cid) + b"\xff\xd9"


def frame_message(rgba: np.ndarray, quality: int = 90) -> bytes:
    return b'{"payload": "data:image/jpeg;base64,' + base64.b64encode(encode(rgba, quality)).rstrip(b"=") + b'"}'
