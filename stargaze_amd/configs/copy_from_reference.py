"""Copy the reference's bundled scene files (its ``configs/``) into this
folder, whole-line comments removed; the data lines are unchanged, so the
reference tests' line substitutions apply to the copies too. The package
loads them as its bundled scenes and the tests as fixtures.

Run in a checkout that has the reference at ``/root/reference``:
``python stargaze_amd/configs/copy_from_reference.py``.
"""
from __future__ import annotations

import pathlib

SOURCE = pathlib.Path("/root/reference/configs")
FILES = [
    "1-halo-volcano.yaml",
    "2-erg-dunes.yaml",
    "3-umbra-valley.yaml",
    "legacy/atacama.yaml",
    "legacy/corona-valley.yaml",
    "legacy/halo.yaml",
    "legacy/median-resonance.yaml",
    "legacy/puzzle.yaml",
    "legacy/solar-system.yaml",
    "legacy/vesper.yaml",
]


def main() -> None:
    out = pathlib.Path(__file__).resolve().parent
    for name in FILES:
        lines = (SOURCE / name).read_text(encoding="utf-8").splitlines(keepends=True)
        data = [line for line in lines if not line.lstrip().startswith("#")]
        (out / pathlib.Path(name).name).write_text("".join(data), encoding="utf-8")


if __name__ == "__main__":
    main()
