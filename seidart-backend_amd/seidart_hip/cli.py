"""`seidartfdtd <input.json> [seismic=true|false] [density_method=...]`:
command-line drop-in of src/fortran/main.f08.

Parses the same JSON keys (src/fortran/seidartio.f08:61-114), the same
key=value flags with the same defaults (main.f08:43-67), pads the grid by
2*cpml (main.f08:69-71), validates density_method (73-81) and dispatches on
Domain.dim and `seismic` (90-119).  Runs in the current directory, like the
Fortran executable.  Exit status: 0 on success; 1 on a blow-up (the
reference exits 0 after printing); 2 on bad usage.

    python -m seidart_hip.cli input.json seismic=false
"""
from __future__ import annotations

import json
import sys

from . import cpmlfdtd
from ._lib import UnstableError


def _get(d, dotted, default=None):
    cur = d
    for part in dotted.split("."):
        if not isinstance(cur, dict) or part not in cur:
            return default
        cur = cur[part]
    return cur


def parse_json(path):
    """Domain and the seismic / electromagnetic sources, seidartio.f08:45-118."""
    with open(path) as f:
        j = json.load(f)
    dom = cpmlfdtd.Domain(
        dim=float(_get(j, "Domain.dim", 2.0)), nx=int(_get(j, "Domain.nx", 0)),
        ny=int(_get(j, "Domain.ny", 1) or 1), nz=int(_get(j, "Domain.nz", 0)),
        dx=float(_get(j, "Domain.dx", 1.0)), dy=float(_get(j, "Domain.dy", 1.0) or 1.0),
        dz=float(_get(j, "Domain.dz", 1.0)), cpml=int(_get(j, "Domain.cpml", 0)),
        nmats=int(_get(j, "Domain.nmats", 0) or 0),
        image_file=str(_get(j, "Domain.image_file", "") or "").strip())

    def source(prefix):
        g = lambda k, d=0.0: _get(j, f"{prefix}.Source.{k}", d)
        return cpmlfdtd.Source(
            dt=float(g("dt")), time_steps=int(g("time_steps", 0)), x=float(g("x")), y=float(g("y")),
            z=float(g("z")), xind=int(g("xind", 1)), yind=int(g("yind", 1)), zind=int(g("zind", 1)),
            source_frequency=float(g("source_frequency")), x_z_rotation=float(g("x-z_rotation")),
            x_y_rotation=float(g("x-y_rotation")), y_z_rotation=float(g("y-z_rotation")),
            amplitude=float(g("amplitude", 1.0)), source_type=str(g("source_type", "")).strip(),
            source_wavelet=str(g("source_wavelet", "")).strip())

    return dom, source("Seismic"), source("Electromagnetic")


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if len(argv) < 1:
        print("Usage: main <input.json> [seismic=<true|false>]")
        return 2
    domain, seis, em = parse_json(argv[0])
    seismic, density_method = True, "arithmetic"
    for arg in argv[1:]:
        if "=" not in arg:
            continue
        key, value = (s.strip() for s in arg.split("=", 1))
        if key == "seismic":
            if value == "true":
                seismic = True
            elif value == "false":
                seismic = False
        elif key == "density_method":
            density_method = value
    domain.nx += 2 * domain.cpml
    domain.ny += 2 * domain.cpml
    domain.nz += 2 * domain.cpml
    if density_method not in ("harmonic", "geometric", "arithmetic", "none"):
        print(f'Error: invalid density_method "{density_method}".')
        print("       Must be one of: harmonic, geometric, arithmetic, none")
        return 2
    try:
        if domain.dim == 3.0:
            if seismic:
                print("Running 3D seismic model with", seis.time_steps, "time steps")
                cpmlfdtd.seismic3(domain, seis, density_method, True)
            else:
                print("Running 3D electromagnetic model with", em.time_steps, "time steps")
                cpmlfdtd.electromag3(domain, em, True)
        elif domain.dim == 2.5:
            if seismic:
                print("Running 2.5D seismic model with", seis.time_steps, "time steps")
                cpmlfdtd.seismic25(domain, seis)
            else:
                print("Running 2.5D electromagnetic model with", em.time_steps, "time steps")
                cpmlfdtd.electromag25(domain, em)
        else:
            # the 2D solvers work on the (x, z) plane
            domain.ny = 1
            if seismic:
                print("Running 2D seismic model with", seis.time_steps, "time steps")
                cpmlfdtd.seismic2(domain, seis, density_method, True)
            else:
                print("Running 2D electromagnetic model with", em.time_steps, "time steps")
                cpmlfdtd.electromag2(domain, em, True)
    except UnstableError as e:
        print("code became unstable and blew up", f"(step {e.step})")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
