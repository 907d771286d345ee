"""Saved-result persistence for the hip SSB backend: exact reuse of a prior
fit or reconstruction when nothing that determines it has changed.

Follows the reference's storage contract (``src/quantem/gpu/ssb/_persistence.py:1-235``,
used by ``SSB.fit``/``SSB.reconstruct`` at ``src/quantem/gpu/ssb/workflow.py:1043-1170``):

* ``result_paths``: ``save_to`` is a directory (``ssb-<operation>.npz`` and
  ``.json`` inside it) or an ``.npz`` path (``.json`` beside it);
* the array file holds ``object_wave`` only; the JSON holds the schema
  number, the generation time, git provenance, the input description, the
  signature and every other result field;
* writes are atomic (temporary file + rename);
* ``load_result`` returns ``None`` on any mismatch -- a different schema or
  signature, a missing or unreadable file -- so the caller recomputes.

The signature names everything that determines the result: this package's
own sources (the HIP kernels included), the data source and shape, the
instrument parameters, the backend settings and the operation's settings.
Saved arrays are loaded with ``allow_pickle=False``.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import tempfile
from dataclasses import fields
from datetime import datetime, timezone
from pathlib import Path
from typing import Callable, Literal

import numpy as np

from .types import SSBResult

SCHEMA = 1
_ARTIFACT_FIELDS = {"object_wave", "reused", "saved_path", "metadata"}
_PACKAGE = Path(__file__).resolve().parents[1]          # quantem/gpu_amd
_SOURCE_SUFFIXES = {".py", ".hip", ".h"}


def software_signature() -> dict[str, object]:
    """Identify the implementation: a digest of the package's sources."""
    digest = hashlib.sha256()
    roots = [_PACKAGE, _PACKAGE.parents[1] / "include"]
    for root in roots:
        if not root.is_dir():
            continue
        for path in sorted(root.rglob("*")):
            if path.is_file() and path.suffix in _SOURCE_SUFFIXES and "__pycache__" not in path.parts:
                digest.update(str(path.relative_to(root.parent)).encode("utf-8"))
                digest.update(path.read_bytes())
    return {"package": "quantem.gpu_amd", "source_sha256": digest.hexdigest()}


def _git_provenance() -> dict[str, object] | None:
    repository = _PACKAGE.parents[1]

    def git(*arguments: str) -> str | None:
        try:
            done = subprocess.run(["git", "-C", str(repository), *arguments], capture_output=True, text=True,
                                  check=False, timeout=10)
        except (OSError, subprocess.SubprocessError):
            return None
        return done.stdout.strip() if done.returncode == 0 else None

    commit = git("rev-parse", "HEAD")
    if commit is None:
        return None
    status = git("status", "--short", "--untracked-files=all")
    return {"commit": commit, "branch": git("branch", "--show-current"), "dirty": bool(status)}


def json_value(value: object) -> object:
    """Deterministic JSON-safe form of one value (sorted keys; arrays as lists)."""
    if isinstance(value, dict):
        return {str(k): json_value(v) for k, v in sorted(value.items(), key=lambda kv: str(kv[0]))}
    if isinstance(value, (list, tuple)):
        return [json_value(v) for v in value]
    if isinstance(value, Path):
        return str(value)
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, np.generic):
        return value.item()
    if value is None or isinstance(value, (bool, int, float, str)):
        return value
    raise TypeError(f"Cannot save SSB metadata value of type {type(value).__name__}.")


def path_signature(path: str | Path) -> dict[str, object]:
    """Describe a source path (sizes, times, inodes) without reading its payload."""
    resolved = Path(path).expanduser().resolve()
    if not resolved.exists():
        return {"path": str(resolved), "missing": True}
    paths = [resolved]
    if resolved.is_dir():
        paths.extend(sorted(p for p in resolved.rglob("*") if p.is_file()))
    files = []
    for item in paths:
        st = item.stat()
        files.append({"path": str(item), "size": int(st.st_size), "mtime_ns": int(st.st_mtime_ns),
                      "ctime_ns": int(st.st_ctime_ns), "device": int(st.st_dev), "inode": int(st.st_ino)})
    return {"path": str(resolved), "files": files}


def result_paths(save_to: str | Path, operation: Literal["fit", "reconstruct"]) -> tuple[Path, Path]:
    """(array path, metadata path) of one saved result."""
    destination = Path(save_to).expanduser().resolve()
    if destination.suffix:
        if destination.suffix.lower() != ".npz":
            raise ValueError("save_to must be a directory or an .npz file, for example "
                             "save_to='results/ssb' or save_to='results/ssb-fit.npz'.")
        return destination, destination.with_suffix(".json")
    return destination / f"ssb-{operation}.npz", destination / f"ssb-{operation}.json"


def _atomic_write(path: Path, write: Callable, mode: str) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    kwargs = {"encoding": "utf-8"} if "b" not in mode else {}
    tmp = tempfile.NamedTemporaryFile(mode=mode, prefix=f".{path.name}.", dir=path.parent, delete=False, **kwargs)
    tmp_path = Path(tmp.name)
    try:
        with tmp:
            write(tmp)
        os.replace(tmp_path, path)
    finally:
        tmp_path.unlink(missing_ok=True)


def _host_array(value) -> np.ndarray:
    if hasattr(value, "detach") and hasattr(value, "cpu"):      # torch tensor
        return value.detach().cpu().numpy()
    return np.asarray(value)


def save_result(result: SSBResult, *, paths: tuple[Path, Path], signature: dict[str, object],
                input_metadata: dict[str, object]) -> SSBResult:
    """Persist one result (atomic .npz + .json) and attach its provenance."""
    array_path, metadata_path = paths
    result_fields = {f.name: getattr(result, f.name) for f in fields(result) if f.name not in _ARTIFACT_FIELDS}
    metadata = {
        "schema": SCHEMA,
        "generated_at_utc": datetime.now(timezone.utc).isoformat(),
        "git": _git_provenance(),
        "input": json_value(input_metadata),
        "signature": json_value(signature),
        "result": json_value(result_fields),
    }
    wave = _host_array(result.object_wave)
    _atomic_write(array_path, lambda f: np.savez(f, object_wave=wave), "wb")
    _atomic_write(metadata_path, lambda f: (json.dump(metadata, f, indent=2), f.write("\n")), "w")
    result.reused = False
    result.saved_path = array_path
    result.metadata = metadata
    return result


def load_result(*, paths: tuple[Path, Path], signature: dict[str, object], device=None) -> SSBResult | None:
    """The saved result whose schema and signature match exactly, else ``None``.

    ``device``: a torch device to place ``object_wave`` on (default: numpy)."""
    array_path, metadata_path = paths
    if not array_path.is_file() or not metadata_path.is_file():
        return None
    try:
        metadata = json.loads(metadata_path.read_text(encoding="utf-8"))
        if metadata.get("schema") != SCHEMA or metadata.get("signature") != json_value(signature):
            return None
        saved = dict(metadata["result"])
        with np.load(array_path, allow_pickle=False) as arrays:
            wave = np.array(arrays["object_wave"])
    except (OSError, ValueError, KeyError, TypeError):
        return None
    if device is not None:
        import torch
        wave = torch.from_numpy(wave).to(device)
    for name in ("scan_sampling_A", "bf_center"):
        if isinstance(saved.get(name), list):
            saved[name] = tuple(saved[name])
    known = {f.name for f in fields(SSBResult)}
    saved = {k: v for k, v in saved.items() if k in known}
    return SSBResult(object_wave=wave, reused=True, saved_path=array_path, metadata=metadata, **saved)


def reuse_or_run(operation: Literal["fit", "reconstruct"], save_to: str | Path, signature: dict[str, object],
                 run: Callable[[], SSBResult], *, input_metadata: dict[str, object], force: bool = False,
                 device=None, verbose: bool = False) -> SSBResult:
    """The reuse rule of ``SSB.fit``/``SSB.reconstruct`` (``workflow.py:1069-1103``):
    load a matching saved result unless ``force``; otherwise run and save."""
    paths = result_paths(save_to, operation)
    if not force:
        reused = load_result(paths=paths, signature=signature, device=device)
        if reused is not None:
            if verbose:
                print(f"Matching SSB result found; loading {paths[0]}")
            return reused
        if verbose and any(p.exists() for p in paths):
            print(f"Saved SSB settings changed; running {operation} end to end")
    result = run()
    result = save_result(result, paths=paths, signature=signature, input_metadata=input_metadata)
    if verbose:
        print(f"SSB result saved to {paths[0]}")
    return result
