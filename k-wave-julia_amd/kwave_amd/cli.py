"""Command-line tool: run / validate / info on k-Wave HDF5 input files.

Mirrors ``KWave.jl/src/cli.jl`` (``KWave.CLI.main``): the same subcommands,
options, messages and exit codes; ``run`` drives the GPU solver.

    python -m kwave_amd.cli run --input sim.h5 --output results.h5 [--pml-size 20] [--pml-alpha 2.0] [-v]
    python -m kwave_amd.cli validate --input sim.h5
    python -m kwave_amd.cli info --input sim.h5
"""
from __future__ import annotations

import os
import sys
from typing import Dict, List, Optional

import numpy as np

USAGE = """KWave.jl — k-space pseudospectral acoustic simulation

Usage:
  kwave run --input <file.h5> --output <results.h5> [options]
  kwave validate --input <file.h5>
  kwave info --input <file.h5>
  kwave help

Run options:
  --input, -i     Input HDF5 file path (required)
  --output, -o    Output HDF5 file path (required)
  --pml-size      PML size in grid points (default: 20)
  --pml-alpha     PML absorption coefficient (default: 2.0)
  --verbose, -v   Verbose output
"""


def main(args: Optional[List[str]] = None) -> int:
    """cli.jl:26-50."""
    args = list(sys.argv[1:] if args is None else args)
    if not args:
        print(USAGE)
        return 1
    cmd, rest = args[0].lower(), args[1:]
    if cmd == "run":
        return _cmd_run(rest)
    if cmd == "validate":
        return _cmd_validate(rest)
    if cmd == "info":
        return _cmd_info(rest)
    if cmd in ("help", "-h", "--help"):
        print(USAGE)
        return 0
    print(f"Unknown command: {cmd}", file=sys.stderr)
    print(USAGE)
    return 1


def _parse_args(args: List[str]) -> Dict[str, str]:
    """cli.jl:71-93: --key value pairs; -i/-o/-v short forms; --verbose takes no value."""
    opts: Dict[str, str] = {}
    i = 0
    while i < len(args):
        arg = args[i]
        if arg.startswith("-"):
            key = arg.lstrip("-")
            key = {"i": "input", "o": "output", "v": "verbose"}.get(key, key)
            if key == "verbose":
                opts[key] = "true"
            elif i < len(args) - 1:
                i += 1
                opts[key] = args[i]
        i += 1
    return opts


def _first(v):
    return np.asarray(v).ravel(order="F")[0]


def _cmd_run(args: List[str]) -> int:
    """cli.jl:95-187: read grid/medium/p0 from the input file, record p_max and p_rms
    everywhere, write every output field and the file attributes."""
    from . import hdf5
    from .model import KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource
    from .solver import kspace_first_order

    opts = _parse_args(args)
    input_file, output_file = opts.get("input"), opts.get("output")
    if input_file is None:
        print("Error: --input is required", file=sys.stderr)
        return 1
    if output_file is None:
        print("Error: --output is required", file=sys.stderr)
        return 1
    if not os.path.isfile(input_file):
        print(f"Error: Input file not found: {input_file}", file=sys.stderr)
        return 1
    try:
        pml_size = int(opts.get("pml-size", "20"))
        pml_alpha = float(opts.get("pml-alpha", "2.0"))
        verbose = opts.get("verbose", "false") == "true"
        if verbose:
            print(f"Loading input file: {input_file}")
        f = hdf5.read_file(input_file)
        ds = {k: v.value for k, v in f.datasets.items()}
        Nx, dx = int(_first(ds["Nx"])), float(_first(ds["dx"]))
        Ny = int(_first(ds["Ny"])) if "Ny" in ds else None
        Nz = int(_first(ds["Nz"])) if "Nz" in ds else None
        dy = float(_first(ds["dy"])) if "dy" in ds else None
        dz = float(_first(ds["dz"])) if "dz" in ds else None
        Nt, dt = int(_first(ds["Nt"])), float(_first(ds["dt"]))
        if verbose:
            print(f"Grid: Nx={Nx}" + (f", Ny={Ny}" if Ny is not None else "") + (f", Nz={Nz}" if Nz is not None else ""))
        if Nz is not None:
            kgrid = KWaveGrid(Nx, dx, Ny, dy, Nz, dz)
        elif Ny is not None:
            kgrid = KWaveGrid(Nx, dx, Ny, dy)
        else:
            kgrid = KWaveGrid(Nx, dx)
        kgrid.dt = dt
        kgrid.Nt = Nt
        kgrid.t_array = np.arange(Nt) * dt
        c0 = ds["c0"] if "c0" in ds else float(_first(ds["c_ref"]))
        rho0 = ds["rho0"] if "rho0" in ds else 1.0
        medium = KWaveMedium(sound_speed=c0, density=rho0)
        source = KWaveSource(p0=ds.get("p0"))
        sensor = KWaveSensor(mask=np.ones(kgrid.grid_size(), dtype=bool), record=["p_max", "p_rms"])
        if verbose:
            print(f"Running simulation (Nt={Nt})...")
        result = kspace_first_order(kgrid, medium, source, sensor, pml_size=pml_size, pml_alpha=pml_alpha)
        if verbose:
            print(f"Writing output to: {output_file}")
        for key in result.keys():
            hdf5.write_matrix(output_file, str(key), result[key])
        hdf5.write_attributes(output_file)
        if verbose:
            print("Done.")
    except Exception as e:  # the reference reports any failure and returns 1
        print(f"Error: {e}", file=sys.stderr)
        return 1
    return 0


def _cmd_validate(args: List[str]) -> int:
    """cli.jl:189-230."""
    from . import hdf5

    opts = _parse_args(args)
    input_file = opts.get("input")
    if input_file is None:
        print("Error: --input is required", file=sys.stderr)
        return 1
    if not os.path.isfile(input_file):
        print(f"Error: File not found: {input_file}", file=sys.stderr)
        return 1
    try:
        ds = {k: v.value for k, v in hdf5.read_file(input_file).datasets.items()}
    except Exception as e:
        print(f"Error reading file: {e}", file=sys.stderr)
        return 1
    missing = [k for k in ("Nx", "dx", "Nt", "dt") if k not in ds]
    if missing:
        print("✗ Invalid: missing fields: " + ", ".join(missing), file=sys.stderr)
        return 1
    print("✓ Valid k-Wave input file")
    print("  Nx = ", _first(ds["Nx"]), sep="")
    if "Ny" in ds:
        print("  Ny = ", _first(ds["Ny"]), sep="")
    if "Nz" in ds:
        print("  Nz = ", _first(ds["Nz"]), sep="")
    print("  Nt = ", _first(ds["Nt"]), sep="")
    print("  dt = ", _first(ds["dt"]), sep="")
    return 0


_JULIA_TYPES = {"float64": "Float64", "float32": "Float32", "int64": "Int64", "int32": "Int32",
                "uint8": "UInt8", "int8": "Int8", "int16": "Int16", "uint16": "UInt16", "uint32": "UInt32",
                "uint64": "UInt64", "float16": "Float16"}


def _cmd_info(args: List[str]) -> int:
    """cli.jl:232-276: file attributes, then each dataset's element type and size."""
    from . import hdf5

    opts = _parse_args(args)
    input_file = opts.get("input")
    if input_file is None:
        print("Error: --input is required", file=sys.stderr)
        return 1
    if not os.path.isfile(input_file):
        print(f"Error: File not found: {input_file}", file=sys.stderr)
        return 1
    try:
        f = hdf5.read_file(input_file)
    except Exception as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    print(f"File: {input_file}")
    if f.attrs:
        print("\nAttributes:")
        for k, v in f.attrs.items():
            print(f"  {k} = {v}")
    print("\nDatasets:")
    for name, d in f.datasets.items():
        v = d.value
        if isinstance(v, str):
            print(f"  {name}: String ()")
            continue
        a = np.asarray(v)
        tp = _JULIA_TYPES.get(a.dtype.name, a.dtype.name)
        shape = "(" + ", ".join(str(s) for s in a.shape) + ("," if a.ndim == 1 else "") + ")"
        print(f"  {name}: {tp} {shape}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
