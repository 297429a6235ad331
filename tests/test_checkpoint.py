"""Checkpoint and resume of the filter state (k3dom_checkpoint_*; no
counterpart in the reference).  A map restored from a checkpoint and given
the same later frames must produce the same results bit for bit as the map
it was taken from, including across a map shift that is still pending when
the checkpoint is taken, for both methods, eagerly and graph-replayed.  A
buffer that is not a whole, intact checkpoint is rejected before any device
work (CPU tests)."""
import ctypes as C
import struct

import numpy as np
import pytest

K3DOM_ERR_INVALID_ARG = 1  # include/k3dom.h
HEADER = 128


def _fnv1a_words(payload):
    h = 0xcbf29ce484222325
    for (w,) in struct.iter_unpack("<Q", payload):
        h = ((h ^ w) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _synthetic_checkpoint(V=256, N=64, method=0):
    """A checkpoint-shaped buffer of the documented layout (version 1)."""
    from k3dom_amd.scene import k3dom_params
    p = k3dom_params(0.2, 8, 4, N, 8)
    pb = bytes(p)
    raw = 4 * (V * 12 + N * 8)
    plen = (raw + 7) & ~7
    payload = np.random.default_rng(1).integers(0, 255, plen, dtype=np.uint8).tobytes()
    hd = struct.pack("<8sIIIiqqIIdffffffQQ", b"K3DOMCK1", 1, HEADER, len(pb), method, V, N, 3, 1, 1.5,
                     0, 0, 0, 0, 0, 0, plen, _fnv1a_words(payload))
    hd = hd + b"\0" * (HEADER - len(hd))
    return bytearray(hd + pb + payload), len(hd) + len(pb)


def _load(lib, buf):
    arr = (C.c_uint8 * len(buf)).from_buffer_copy(bytes(buf)) if len(buf) else (C.c_uint8 * 1)()
    h = C.c_void_p()
    rc = lib.k3dom_checkpoint_load(arr, len(buf), C.byref(h))
    return rc, h


def test_load_rejects_damaged_buffers():
    from k3dom_amd import _lib
    lib = _lib.load()
    good, at = _synthetic_checkpoint()
    cases = {"empty": bytearray(), "short": good[:100], "truncated": good[:-8], "extra": good + b"\0" * 8}
    bad_magic = bytearray(good); bad_magic[0] ^= 1
    bad_version = bytearray(good); bad_version[8] = 2
    bad_params = bytearray(good); bad_params[16] ^= 4
    flipped = bytearray(good); flipped[at + 100] ^= 0x10
    bad_size = bytearray(good); bad_size[24] ^= 1   # V no longer matches the payload size
    cases.update(magic=bad_magic, version=bad_version, params_bytes=bad_params, checksum=flipped, sizes=bad_size)
    for name, buf in cases.items():
        rc, h = _load(lib, buf)
        assert rc == K3DOM_ERR_INVALID_ARG and not h.value, name
        assert lib.k3dom_last_error(None), name
    assert lib.k3dom_checkpoint_load(None, 0, C.byref(C.c_void_p())) == K3DOM_ERR_INVALID_ARG


def _run(maps, frame):
    """Runs one frame on every map; True when it was a normal frame."""
    t, pose, xs, ys, zs = frame
    normal = maps[0].info().updated_time > 0
    for m in maps:
        m.updatePose(*pose)
        m.updateGrid(t, xs, ys, zs)
    return normal


def _compare(a, b, what, normal=True):
    from tests._parity import compare_state
    compare_state(a, b, normal_frame=normal, what=what)
    ia, ib = a.info(), b.info()
    assert ia.iteration == ib.iteration and ia.updated_time == ib.updated_time
    assert list(ia.center_pos) == list(ib.center_pos) and list(ia.sensor_pos) == list(ib.sensor_pos)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["k3dom", "dsphdmib"])
def test_resume_continues_bit_exact(gpu, method, tmp_path):
    """Checkpoint after a shifting updatePose (shift still pending), restore
    from a file, continue both maps over the same frames (the restored one
    graph-replayed from its second frame on); the restored map also equals
    the serial restatement run through every frame without interruption."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    from oracle.orc import OracleDOM
    from tests._parity import compare_state
    w = small_workload(ego_vx=4.0, frames=11, seed=3, method=method)
    sc = Scene(w.scene)
    frames = [sc.frame(f) for f in range(w.frames)]
    a = DOM(w.params, method=method)
    o = OracleDOM(w.params, method=method)  # the serial restatement, never interrupted
    k = None
    for f in range(w.frames):
        t, pose, xs, ys, zs = frames[f]
        c0 = a.info().center_pos[0]
        a.updatePose(*pose)
        o.updatePose(*pose)
        if f >= 3 and a.info().center_pos[0] != c0:
            k = f
            break
        a.updateGrid(t, xs, ys, zs)
        o.updateGrid(t, xs, ys, zs)
    assert k is not None, "no map shift in the first frames"
    print(f"  {method}: checkpoint after the shifting pose of frame {k}", flush=True)
    path = str(tmp_path / "map.k3ck")
    a.save(path)
    b = DOM.load(path)
    assert b.method == method and b.params.particle_count == w.params.particle_count
    t, pose, xs, ys, zs = frames[k]
    for m in (a, b, o):
        m.updateGrid(t, xs, ys, zs)
    _compare(a, b, f"frame {k} after resume:")
    compare_state(b, o, normal_frame=True, what=f"frame {k} after resume vs restatement:")
    b.set_graph(True)
    for f in range(k + 1, w.frames):
        _run((a, b, o), frames[f])
        _compare(a, b, f"frame {f} after resume:")
        compare_state(b, o, normal_frame=True, what=f"frame {f} after resume vs restatement:")


@pytest.mark.gpu
def test_checkpoint_of_a_checkpoint_and_first_frame_state(gpu):
    """A map restored before its first frame initialises like a new map; a
    checkpoint of a restored map equals the original checkpoint."""
    from k3dom_amd.dom import DOM
    from k3dom_amd.scene import Scene, small_workload
    w = small_workload(frames=4, seed=5)
    sc = Scene(w.scene)
    frames = [sc.frame(f) for f in range(w.frames)]
    a = DOM(w.params)
    b = DOM.from_checkpoint(a.checkpoint())
    for f in range(2):
        normal = _run((a, b), frames[f])
        _compare(a, b, f"frame {f} from a fresh checkpoint:", normal)
    data = a.checkpoint()
    c = DOM.from_checkpoint(data)
    assert c.checkpoint() == data
    for f in range(2, w.frames):
        _run((a, c), frames[f])
        _compare(a, c, f"frame {f}:")
