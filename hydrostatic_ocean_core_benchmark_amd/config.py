"""TOML configuration tree: five files deep-merged in a fixed order, dotted-key
lookup, dotted-key overrides and a stable hash for the run manifest.

Behaviour follows the reference loader (libs/utils/config.py:22-98) so that a
config directory written for the reference loads unchanged, and `--set
KEY=VALUE` parsing follows main.py:335-349 (int, then float, then
true/false, else the raw string).
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path
from typing import Any, Iterable

try:                                   # Python >= 3.11
    import tomllib
except ModuleNotFoundError:            # Python 3.10
    import tomli as tomllib            # type: ignore[no-redef]

# Merge order; later files override earlier keys (libs/utils/config.py:22).
CONFIG_FILES = ("settings.toml", "grid.toml", "schemes.toml", "cases.toml",
                "backends.toml")

DEFAULT_CONFIG_DIR = Path(__file__).resolve().parents[1] / "config"


class ConfigError(KeyError):
    """A required key is missing or is not a table."""


_MISSING = object()


class Config:
    """Read-only view of the merged configuration mapping."""

    def __init__(self, data: dict[str, Any], source_dir: Path) -> None:
        self._data = data
        self.source_dir = Path(source_dir)

    def get(self, dotted: str, default: Any = _MISSING) -> Any:
        node: Any = self._data
        for part in dotted.split("."):
            if not isinstance(node, dict) or part not in node:
                if default is _MISSING:
                    raise ConfigError(f"missing config key '{dotted}' (at '{part}')")
                return default
            node = node[part]
        return node

    def section(self, dotted: str) -> dict[str, Any]:
        node = self.get(dotted)
        if not isinstance(node, dict):
            raise ConfigError(f"config key '{dotted}' is not a table")
        return dict(node)

    def as_dict(self) -> dict[str, Any]:
        return json.loads(json.dumps(self._data))

    def sha256(self) -> str:
        blob = json.dumps(self._data, sort_keys=True, separators=(",", ":"))
        return hashlib.sha256(blob.encode("utf-8")).hexdigest()

    def with_overrides(self, overrides: dict[str, Any]) -> "Config":
        data = self.as_dict()
        for dotted, value in overrides.items():
            node = data
            *head, leaf = dotted.split(".")
            for part in head:
                node = node.setdefault(part, {})
            node[leaf] = value
        return Config(data, self.source_dir)


def _merge_into(base: dict[str, Any], extra: dict[str, Any]) -> dict[str, Any]:
    for key, value in extra.items():
        if isinstance(value, dict) and isinstance(base.get(key), dict):
            _merge_into(base[key], value)
        else:
            base[key] = value
    return base


def load_config(config_dir: Path | str | None = None) -> Config:
    directory = Path(config_dir) if config_dir is not None else DEFAULT_CONFIG_DIR
    if not directory.is_dir():
        raise FileNotFoundError(f"config directory not found: {directory}")
    merged: dict[str, Any] = {}
    for name in CONFIG_FILES:
        path = directory / name
        if not path.exists():
            raise FileNotFoundError(f"required config file missing: {path}")
        with path.open("rb") as fh:
            _merge_into(merged, tomllib.load(fh))
    return Config(merged, directory)


def parse_value(raw: str) -> Any:
    """`--set` value inference: int, float, true/false, else string."""
    try:
        return int(raw)
    except ValueError:
        pass
    try:
        return float(raw)
    except ValueError:
        return {"true": True, "false": False}.get(raw.lower(), raw)


def apply_overrides(config: Config, overrides: Iterable[str]) -> Config:
    parsed: dict[str, Any] = {}
    for item in overrides:
        if "=" not in item:
            raise SystemExit(f"--set expects KEY=VALUE, got '{item}'")
        key, raw = item.split("=", 1)
        parsed[key] = parse_value(raw)
    return config.with_overrides(parsed) if parsed else config
