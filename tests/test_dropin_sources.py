"""Source-type dispatch of the file-based drop-in (no GPU): which source
types select a plane wave and of which kind (plane_wave_source.f08:20-35 for
the elastic solvers, `source_type == 'pw'` for electromag2/3,
cpmlfdtd.f08:2436, 3523), the wavelet names (plane_wave_source.f08:150-160),
and electromag25's refusal of 'pw'."""
import pytest

from seidart_hip import cpmlfdtd


class _Recorder:
    def __init__(self):
        self.calls = []

    def set_plane_wave(self, kind, angles, f0, amplitude, wavelet=0):
        self.calls.append((kind, tuple(angles), f0, amplitude, wavelet))


def _src(st, wavelet="gaus2"):
    return cpmlfdtd.Source(dt=1e-4, time_steps=10, source_type=st, source_wavelet=wavelet, source_frequency=50.0,
                           x_z_rotation=10.0, x_y_rotation=20.0, y_z_rotation=30.0, amplitude=2.0)


@pytest.mark.parametrize("st,kind", [("pw", "s"), ("pws", "s"), ("pw_s", "s"), ("pw-s", "s"), ("  pw-s ", "s"),
                                     ("pwp", "p"), ("pw_p", "p"), ("pw-p", "p"), (" pwp", "p"),
                                     ("ac", None), ("tensor", None), ("", None), ("pw s", None), ("PW", None)])
def test_elastic_source_types(st, kind):
    r = _Recorder()
    assert cpmlfdtd._elastic_plane_wave(r, _src(st)) == (kind is not None)
    if kind is None:
        assert r.calls == []
    else:
        assert r.calls == [(kind, (10.0, 20.0, 30.0), 50.0, 2.0, 2)]


@pytest.mark.parametrize("st,pw", [("pw", True), ("pw  ", True), (" pw", False), ("pws", False), ("pwp", False),
                                   ("ac", False)])
def test_em_source_types(st, pw):
    r = _Recorder()
    assert cpmlfdtd._em_plane_wave(r, _src(st)) == pw
    assert r.calls == ([("s", (10.0, 20.0, 30.0), 50.0, 2.0, 2)] if pw else [])


@pytest.mark.parametrize("name,code", [("gaus0", 0), ("gaus1", 1), ("gaus2", 2), ("gaus2  ", 2), ("ricker", 0),
                                       ("", 0), (" gaus1", 0)])
def test_wavelet_names(name, code):
    assert cpmlfdtd._wavelet(_src("pw", name)) == code


def test_electromag25_rejects_plane_wave(tmp_path):
    dom = cpmlfdtd.Domain(dim=2.5, nx=20, ny=1, nz=20, dx=0.1, dy=0.1, dz=0.1, cpml=4)
    with pytest.raises(ValueError, match="electromag25"):
        cpmlfdtd.electromag25(dom, _src("pw"), workdir=str(tmp_path))
