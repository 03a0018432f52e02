"""The temperature operator (calculate_temperature_operator.cu:27-78) and the visual-stream
packets (output/server_writer.cu:79-252, datastruct.h:4-72) on the GPU.

- temperature grid: theta(x, y, z) = sum_i S_i(z) C2R(theta_i) vs the FP64 numpy sum (2e-6 of
  the field's maximum in FP32);
- packets: 2048-byte packets, packet 0 of type 0x491C with the simulation header, data packets
  of type 0x08B1 holding the image sampled x outer / y inner in chunks of 506 floats; the
  particle image counts every particle once."""
import struct

import numpy as np
import pytest

from oracle import planes as PL
from tests.helpers import random_state

pytestmark = pytest.mark.gpu

CFG = (64, 32, 8, 4.0, 2.0, 2e4, 0.7, 1e-3)


def _state():
    return random_state(64, 32, 8, seed=71, amp=1e-2)


def _want_temperature(st, z):
    S = np.array([np.float32(np.sin(np.pi * (i + 1) * (np.float32(z) + 0.5))) for i in range(8)], np.float64)
    return sum(S[i] * PL.c2r(st[0][i], 64, 32) for i in range(8))


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_temperature_grid_vs_oracle(precision):
    print("case: temperature grid", precision, flush=True)
    from sdc_amd import Stepper
    st = _state()
    z = np.array([0.0, -0.3, 0.45])
    with Stepper(*CFG, precision=precision) as S:
        S.set_state(*st)
        got = S.temperature_grid(z)
    for k, zz in enumerate(z):
        want = _want_temperature(st, zz)
        assert np.abs(got[k] - want).max() < 2e-6 * np.abs(want).max()


def _parse(raw):
    assert len(raw) % 2048 == 0
    packets = [raw[i:i + 2048] for i in range(0, len(raw), 2048)]
    heads = [struct.unpack("<HHIII", p[:16]) for p in packets]
    return packets, heads


def test_stream_packets_temperature_and_particles():
    print("case: stream packets", flush=True)
    from sdc_amd import Stepper, Tracer
    st = _state()
    with Stepper(*CFG, precision="f32") as S:
        S.set_state(*st)
        raw = S.stream_packets("temperature", res=0, Nt=17, eta=0.0, t_final=2.5)
        T = Tracer(S, 300, buffer_size=8, mode="rk4_trilinear", seed=3)
        T.init_positions("random")
        T.step_particles(1e-3)
        T.step_particles(1e-3)
        rawp = S.stream_packets("particles", res=64, Nt=18, tracer=T)
        xs, ys, _ = T.get_particle_positions(clear=False)
        T.close()
    packets, heads = _parse(raw)
    n = 1 + (64 * 32 + 505) // 506
    assert len(packets) == n
    assert heads[0] == (0x491C, n, 17, 64, 32) and all(h == (0x08B1, n, 17, 64, 32) for h in heads[1:])
    hdr = struct.unpack("<10f3H", packets[0][16:16 + 46])
    assert abs(hdr[0] - 1.1) < 1e-6 and hdr[1] == 1.0 and hdr[2] == 2.5
    assert abs(hdr[3] - 1e-3) < 1e-9 and hdr[4] == 2e4 and abs(hdr[5] - 0.7) < 1e-6
    assert hdr[7:10] == (4.0, 2.0, 1.0) and hdr[10:13] == (64, 32, 8)
    descr = packets[0][16 + 46 + 32:16 + 46 + 64].split(b"\0")[0].decode()
    assert descr == "Temperature field at z = 0"
    img = np.zeros(64 * 32, np.float32)
    for p in packets[1:]:
        i0, i1 = struct.unpack("<ii", p[16:24])
        img[i0:i1] = np.frombuffer(p[24:24 + 4 * (i1 - i0)], np.float32)
    want = _want_temperature(st, 0.0)  # [y][x]; packets run x outer, y inner
    assert np.abs(img.reshape(64, 32) - want.T).max() < 2e-6 * np.abs(want).max()
    packets, heads = _parse(rawp)
    assert heads[0][3:] == (64, 64)
    img = np.zeros(64 * 64, np.float32)
    for p in packets[1:]:
        i0, i1 = struct.unpack("<ii", p[16:24])
        img[i0:i1] = np.frombuffer(p[24:24 + 4 * (i1 - i0)], np.float32)
    assert img.sum() == 300
    # the last read-out slot (noClear) is the image's time
    x, y = xs[-1].astype(np.float64), ys[-1].astype(np.float64)
    ix = np.floor(64 * x / 4.0).astype(int) % 64
    iy = np.floor(64 * y / 2.0).astype(int) % 64
    want = np.zeros((64, 64))
    np.add.at(want, (ix, iy), 1)
    assert np.array_equal(img.reshape(64, 64), want)
