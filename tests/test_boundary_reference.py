"""The drop-in boundary against the reference's own interface headers (CPU;
skipped when the reference tree is absent, as on GPU hosts).

* qusim_amd/host/hip_device.cpp compiles with the reference's src/Device.h
  and src/FourierTransform.h in place of the restated interfaces, and
  HipDevice is a concrete Device there (every pure virtual overridden);
* the restated Device (qusim_amd/host/device_iface.h) declares the same
  virtual methods in the same order as src/Device.h:37-174, so the vtable of
  libqsim_host.so lines up with what reference code calls."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
HOST = os.path.join(ROOT, "qusim_amd", "host")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "src", "Device.h")),
                                reason="reference tree absent")


def _virtuals(path, struct="Device"):
    text = open(path).read()
    m = re.search(r"struct\s+%s\s*\{" % struct, text)
    body = text[m.end():]
    body = body[:body.index("\n};")]
    body = re.sub(r"//[^\n]*", "", body)
    decls = re.findall(r"virtual\s+([^;]*?)\s*=\s*0\s*;", body, re.S)
    out = []
    for d in decls:
        d = re.sub(r"\s+", " ", d).strip()
        d = re.sub(r"\s*([(),*&])\s*", r"\1", d)
        # parameter names do not matter for the vtable: keep types only
        name, _, args = d.partition("(")
        args = args.rsplit(")", 1)[0]
        types = []
        for a in filter(None, args.split(",")):
            a = a.strip()
            a = re.sub(r"\b[A-Za-z_]\w*$", "", a) if re.search(r"[\w*&]\s*\w+$", a) and " " in a else a
            types.append(re.sub(r"\s+", "", a))
        out.append((re.sub(r"\s+", "", name), tuple(types)))
    return out


def test_restated_device_matches_reference_vtable_order():
    ref = _virtuals(os.path.join(REF, "src", "Device.h"))
    ours = _virtuals(os.path.join(HOST, "device_iface.h"))
    assert len(ref) == 64
    assert [r[0] for r in ours] == [r[0] for r in ref]
    assert ours == ref


def test_hip_device_compiles_against_reference_headers(tmp_path):
    probe = tmp_path / "probe.cpp"
    probe.write_text('#include "hip_device.cpp"\n'
                     'Device *qsim_probe_make() { return new qsim_host::HipDevice(); }\n')
    cmd = ["g++", "-fsyntax-only", "-std=c++17", "-D__declspec(x)=", "-DQSIM_FFT_IFACE",
           "-I" + os.path.join(REF, "src"), "-I" + os.path.join(REF, "include"),
           "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
           "-include", os.path.join(REF, "src", "Device.h"),
           "-include", os.path.join(REF, "src", "FourierTransform.h"), str(probe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
