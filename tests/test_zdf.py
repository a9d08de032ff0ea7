"""ZDF diagnostic files written by the *_report functions: byte-identical
to the reference's files for the same state (em2d/zdf.c writer,
em2d/emf.c:368-480, em2d/current.c:195-281, em2d/particles.c:1339-1740)."""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from conftest import HERE, ROOT
from em2d_driver import Em2dRun
import ref_worker

LIB = os.path.join(ROOT, "zpic_amd", "libzpic_em2d.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_em2d.so")


def files(root):
    out = {}
    for d, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(d, f)
            out[os.path.relpath(p, root)] = open(p, "rb").read()
    return out


def read_zdf(blob: bytes):
    """Minimal reader: returns {record name: payload-dict or array}."""
    assert blob[:4] == b"ZDF1"
    pos, out = 4, {}

    def u32():
        nonlocal pos
        v = struct.unpack_from("<I", blob, pos)[0]
        pos += 4
        return v

    def s():
        nonlocal pos
        n = u32()
        v = blob[pos:pos + n].decode()
        pos += (n + 3) // 4 * 4
        return v

    while pos < len(blob):
        rid = u32()
        name = s()
        length = struct.unpack_from("<Q", blob, pos)[0]
        pos += 8
        start = pos
        if rid == 0x00100002:
            _, dtype, ndims = struct.unpack_from("<IiI", blob, pos)
            count = struct.unpack_from("<%dQ" % ndims, blob, pos + 12)
            data = np.frombuffer(blob, dtype="<f4", count=int(np.prod(count)), offset=pos + 12 + 8 * ndims)
            out[name] = data.reshape(tuple(reversed(count)))
        pos = start + length
    return out


@pytest.mark.skipif(not (os.path.exists(LIB) and os.path.exists(REF)), reason="libraries not built")
def test_reports_are_byte_identical_to_the_reference(tmp_path):
    """At the initial state (host-side initialisation is bit-exact) every
    field, current, particle and phase-space report matches the
    reference's file byte for byte."""
    ref_dir = tmp_path / "ref"
    out = subprocess.run([sys.executable, os.path.join(HERE, "ref_worker.py"), "report", str(ref_dir)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-400:]
    lib = C.CDLL(LIB)
    run = Em2dRun(lib, ref_worker.report_case())
    try:
        ref_worker.write_reports(lib, run, str(tmp_path / "gpu"))
    finally:
        run.close()
    a, b = files(ref_dir), files(tmp_path / "gpu")
    assert sorted(a) == sorted(b)
    assert len(a) == 4 * 3 + 3 + 3
    for k in a:
        assert a[k] == b[k], k
    # the reader sees the grids the reports were made from
    e1 = read_zdf(a["EMF/E1-000000.zdf"])["E1"]
    assert e1.shape == (16, 120) and np.abs(e1).max() > 0


def split_zdf(blob: bytes):
    """(bytes outside dataset payloads, list of payload arrays)"""
    assert blob[:4] == b"ZDF1"
    pos, meta, data = 4, bytearray(blob[:4]), []
    while pos < len(blob):
        rid = struct.unpack_from("<I", blob, pos)[0]
        n = struct.unpack_from("<I", blob, pos + 4)[0]
        hdr_end = pos + 8 + (n + 3) // 4 * 4 + 8
        length = struct.unpack_from("<Q", blob, hdr_end - 8)[0]
        if rid == 0x00100002:
            ndims = struct.unpack_from("<I", blob, hdr_end + 8)[0]
            body = hdr_end + 12 + 8 * ndims
            meta += blob[pos:body]
            data.append(np.frombuffer(blob[body:hdr_end + length], dtype="<f4"))
        else:
            meta += blob[pos:hdr_end + length]
        pos = hdr_end + length
    return bytes(meta), data


@pytest.mark.gpu
def test_reports_after_steps_match_the_reference(tmp_path):
    """After 12 steps on the GPU: the same files as the reference's, with
    identical headers; grids within FP32 rounding of the current sum
    (1e-5 of the largest value), particle quantities equal as sets (the
    files list particles in tile order), charge within 1e-5."""
    env = dict(os.environ, ZPIC_REPORT_CHARGE="1")
    out = subprocess.run([sys.executable, os.path.join(HERE, "ref_worker.py"), "report", str(tmp_path / "ref"), "12"],
                         capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr[-400:]
    lib = C.CDLL(LIB)
    run = Em2dRun(lib, ref_worker.report_case(), sync=lambda r: lib.sim_sync_host(C.byref(r.sim)))
    try:
        run.step(12)
        os.environ["ZPIC_REPORT_CHARGE"] = "1"
        try:
            ref_worker.write_reports(lib, run, str(tmp_path / "gpu"))
        finally:
            del os.environ["ZPIC_REPORT_CHARGE"]
    finally:
        run.close()
    a, b = files(tmp_path / "ref"), files(tmp_path / "gpu")
    assert sorted(a) == sorted(b)
    for k in a:
        ma, da = split_zdf(a[k])
        mb, db = split_zdf(b[k])
        assert ma == mb, k
        if k.startswith("PARTICLES"):
            for x, y in zip(da, db):
                assert np.allclose(np.sort(x), np.sort(y), rtol=0, atol=1e-4), k
        else:
            for x, y in zip(da, db):
                scale = max(float(np.abs(x).max()), 1e-30)
                assert np.abs(x - y).max() <= 1e-5 * scale, k
