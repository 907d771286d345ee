"""Saved SSB results: reuse a prior fit or reconstruction when nothing that
determines it has changed.

On-disk format (the interchange format of the reference's saved results,
``src/quantem/gpu/ssb/_persistence.py:1-235``, used by ``SSB.fit`` and
``SSB.reconstruct`` at ``src/quantem/gpu/ssb/workflow.py:933-996``,
``1043-1170``), kept so saved files read the same way:

* ``save_to`` names a directory (``ssb-<operation>.npz`` and
  ``ssb-<operation>.json`` inside it) or an ``.npz`` file (its ``.json``
  beside it);
* the ``.npz`` holds one array, ``object_wave``;
* the ``.json`` holds ``schema`` (1), ``generated_at_utc``, ``git``,
  ``input``, ``signature`` and ``result`` (every other ``SSBResult`` field).

This module's own design: a ``ResultStore`` per (location, operation)
publishes both files by writing complete bytes to an exclusive temporary
file in the target directory, flushing it to disk and renaming it over the
target, so a reader never sees a partial file.  Metadata is normalised by
one JSON round trip with sorted keys (numpy values, paths and tuples
encoded on the way), so a signature compares equal exactly when its
canonical JSON text does.  The software identity is a digest over per-file
digests of this package's sources (kernels included); the git state comes
from one ``git status --porcelain=v2 --branch`` call.  Any problem reading
a saved pair (missing, unreadable, a different schema or signature) means
"no reusable result": the caller recomputes.  Arrays load with
``allow_pickle=False``.
"""
from __future__ import annotations

import hashlib
import io
import json
import os
import subprocess
from dataclasses import dataclass, fields
from datetime import datetime, timezone
from pathlib import Path
from typing import Callable, Literal

import numpy as np

from .types import SSBResult

SCHEMA = 1
Operation = Literal["fit", "reconstruct"]

_PACKAGE_ROOT = Path(__file__).resolve().parents[1]            # quantem/gpu_amd
_REPO_ROOT = _PACKAGE_ROOT.parents[1]
_HASHED_SUFFIXES = (".py", ".hip", ".h")
# fields of SSBResult that describe the stored artefact rather than the result
_NOT_IN_METADATA = frozenset({"object_wave", "reused", "saved_path", "metadata"})
# JSON arrays that SSBResult holds as tuples
_TUPLE_FIELDS = ("scan_sampling_A", "bf_center")


# --------------------------------------------------------------- encoding
def _encode_extra(value):
    """``json.dumps`` hook for the non-JSON types metadata may hold."""
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, np.generic):
        return value.item()
    if isinstance(value, os.PathLike):
        return os.fspath(value)
    raise TypeError(f"Cannot save SSB metadata value of type {type(value).__name__}.")


def _stringify_keys(value):
    """Dict keys as strings (JSON's rule), recursively, before sorting."""
    if isinstance(value, dict):
        return {str(k): _stringify_keys(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_stringify_keys(v) for v in value]
    return value


def canonical_text(value: object) -> str:
    """Canonical JSON text of ``value``: sorted keys, no insignificant space."""
    return json.dumps(_stringify_keys(value), sort_keys=True, separators=(",", ":"), default=_encode_extra)


def json_value(value: object) -> object:
    """The JSON-safe form of ``value`` (what saving and reloading gives):
    dicts with sorted string keys, tuples and arrays as lists, numpy
    scalars as Python numbers, paths as strings.  Any other type raises
    ``TypeError``."""
    return json.loads(canonical_text(value))


# ------------------------------------------------------------- provenance
def _file_digest(path: Path) -> str:
    h = hashlib.sha256()
    with path.open("rb") as fh:
        for block in iter(lambda: fh.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


def software_signature() -> dict[str, object]:
    """Identity of the implementation: one digest over the sorted
    ``relative path  file digest`` lines of the package's Python and HIP
    sources and of the public C headers."""
    lines = []
    for root in (_PACKAGE_ROOT, _REPO_ROOT / "include"):
        if not root.is_dir():
            continue
        for path in root.rglob("*"):
            if path.suffix in _HASHED_SUFFIXES and path.is_file() and "__pycache__" not in path.parts:
                lines.append(f"{path.relative_to(_REPO_ROOT).as_posix()}  {_file_digest(path)}")
    manifest = "\n".join(sorted(lines)).encode("utf-8")
    return {"package": "quantem.gpu_amd", "source_sha256": hashlib.sha256(manifest).hexdigest()}


def _git_state() -> dict[str, object] | None:
    """{commit, branch, dirty} of the checkout this package runs from, or
    ``None`` outside a git checkout (or without git)."""
    try:
        done = subprocess.run(["git", "-C", str(_REPO_ROOT), "status", "--porcelain=v2", "--branch",
                               "--untracked-files=all"], capture_output=True, text=True, timeout=10)
    except (OSError, subprocess.SubprocessError):
        return None
    if done.returncode != 0:
        return None
    header: dict[str, str] = {}
    dirty = False
    for line in done.stdout.splitlines():
        if line.startswith("# branch."):
            key, _, val = line[len("# branch."):].partition(" ")
            header[key] = val.strip()
        elif line.strip():
            dirty = True
    commit = header.get("oid")
    if not commit or commit == "(initial)":
        return None
    branch = header.get("head")
    return {"commit": commit, "branch": None if branch in (None, "(detached)") else branch, "dirty": dirty}


def path_signature(path: str | Path) -> dict[str, object]:
    """What identifies a detector source without reading it: for the path
    (and, for a directory, every file below it, in sorted order) its size,
    modification and change times, device and inode."""
    root = Path(path).expanduser().resolve()
    if not root.exists():
        return {"path": str(root), "missing": True}
    members = [root]
    if root.is_dir():
        for dirpath, dirnames, filenames in os.walk(root):
            dirnames.sort()
            members.extend(Path(dirpath) / name for name in sorted(filenames))
        members = [root] + sorted(members[1:])

    def describe(p: Path) -> dict[str, object]:
        st = p.stat()
        return {"path": str(p), "size": st.st_size, "mtime_ns": st.st_mtime_ns, "ctime_ns": st.st_ctime_ns,
                "device": st.st_dev, "inode": st.st_ino}
    return {"path": str(root), "files": [describe(p) for p in members]}


# ----------------------------------------------------------------- storage
def result_paths(save_to: str | Path, operation: Operation) -> tuple[Path, Path]:
    """(array file, metadata file) of one saved result."""
    target = Path(save_to).expanduser().resolve()
    if not target.suffix:
        stem = target / f"ssb-{operation}"
        return stem.with_suffix(".npz"), stem.with_suffix(".json")
    if target.suffix.lower() == ".npz":
        return target, target.with_suffix(".json")
    raise ValueError("save_to must be a directory or an .npz file, for example "
                     "save_to='results/ssb' or save_to='results/ssb-fit.npz'.")


def _publish(target: Path, payload: bytes) -> None:
    """Replace ``target`` with ``payload`` in one rename: the bytes go to an
    exclusive temporary file in the same directory, are flushed to disk,
    and the file is renamed over the target."""
    target.parent.mkdir(parents=True, exist_ok=True)
    for attempt in range(100):
        tmp = target.parent / f".{target.name}.{os.getpid()}.{attempt}.part"
        try:
            fd = os.open(tmp, os.O_WRONLY | os.O_CREAT | os.O_EXCL, 0o644)
        except FileExistsError:
            continue
        try:
            with os.fdopen(fd, "wb") as fh:
                fh.write(payload)
                fh.flush()
                os.fsync(fh.fileno())
            os.replace(tmp, target)
        finally:
            if tmp.exists():
                tmp.unlink()
        return
    raise OSError(f"no free temporary name beside {target}")


def _to_host(array) -> np.ndarray:
    if hasattr(array, "detach"):                      # torch tensor, possibly on the GPU
        array = array.detach().cpu().numpy()
    return np.asarray(array)


@dataclass(frozen=True)
class ResultStore:
    """The saved result of one operation at one location."""

    array_path: Path
    metadata_path: Path

    @classmethod
    def at(cls, save_to: str | Path, operation: Operation) -> "ResultStore":
        return cls(*result_paths(save_to, operation))

    @property
    def paths(self) -> tuple[Path, Path]:
        return self.array_path, self.metadata_path

    def exists(self) -> bool:
        return self.array_path.exists() or self.metadata_path.exists()

    def save(self, result: SSBResult, signature: dict, input_metadata: dict) -> SSBResult:
        """Write both files and attach the provenance to ``result``."""
        payload = {f.name: getattr(result, f.name) for f in fields(result) if f.name not in _NOT_IN_METADATA}
        metadata = {
            "schema": SCHEMA,
            "generated_at_utc": datetime.now(timezone.utc).isoformat(),
            "git": _git_state(),
            "input": json_value(input_metadata),
            "signature": json_value(signature),
            "result": json_value(payload),
        }
        buf = io.BytesIO()
        np.savez(buf, object_wave=_to_host(result.object_wave))
        _publish(self.array_path, buf.getvalue())
        _publish(self.metadata_path, (json.dumps(metadata, indent=2) + "\n").encode("utf-8"))
        result.reused = False
        result.saved_path = self.array_path
        result.metadata = metadata
        return result

    def load(self, signature: dict, device=None) -> SSBResult | None:
        """The stored result if its schema and signature match ``signature``
        exactly; ``None`` otherwise.  ``device``: torch device for
        ``object_wave`` (default: a NumPy array)."""
        try:
            metadata = json.loads(self.metadata_path.read_bytes())
            if not isinstance(metadata, dict) or metadata.get("schema") != SCHEMA:
                return None
            if canonical_text(metadata.get("signature")) != canonical_text(signature):
                return None
            stored = dict(metadata["result"])
            with np.load(self.array_path, allow_pickle=False) as arrays:
                wave = np.array(arrays["object_wave"])
        except (OSError, ValueError, KeyError, TypeError):
            return None
        known = {f.name for f in fields(SSBResult)} - _NOT_IN_METADATA
        kwargs = {k: v for k, v in stored.items() if k in known}
        for name in _TUPLE_FIELDS:
            if isinstance(kwargs.get(name), list):
                kwargs[name] = tuple(kwargs[name])
        if device is not None:
            import torch
            wave = torch.from_numpy(wave).to(device)
        return SSBResult(object_wave=wave, reused=True, saved_path=self.array_path, metadata=metadata, **kwargs)


def save_result(result: SSBResult, *, paths: tuple[Path, Path], signature: dict,
                input_metadata: dict) -> SSBResult:
    """Persist ``result`` at ``paths`` (see ``result_paths``)."""
    return ResultStore(*paths).save(result, signature, input_metadata)


def load_result(*, paths: tuple[Path, Path], signature: dict, device=None) -> SSBResult | None:
    """The result saved at ``paths`` if it matches ``signature``, else ``None``."""
    return ResultStore(*paths).load(signature, device)


def reuse_or_run(operation: Operation, save_to: str | Path, signature: dict, run: Callable[[], SSBResult], *,
                 input_metadata: dict, force: bool = False, device=None, verbose: bool = False) -> SSBResult:
    """Reuse rule of ``SSB.fit``/``SSB.reconstruct``
    (``src/quantem/gpu/ssb/workflow.py:1069-1103``): a matching saved result
    is returned unless ``force``; otherwise ``run()`` and save."""
    store = ResultStore.at(save_to, operation)
    if not force:
        hit = store.load(signature, device)
        if hit is not None:
            if verbose:
                print(f"Matching SSB result found; loading {store.array_path}")
            return hit
        if verbose and store.exists():
            print(f"Saved SSB settings changed; running {operation} end to end")
    result = store.save(run(), signature, input_metadata)
    if verbose:
        print(f"SSB result saved to {store.array_path}")
    return result
