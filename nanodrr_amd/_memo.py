"""Tensors derived from other tensors, kept until their source changes.

An entry holds a strong reference to the source tensor it was made from and
the source's version counter at that time. It is valid while the source is
the same Python object (identity, not address: a freed tensor's memory can
come back at the same address with the same version) at the same version,
and the extra key (layout, sample count, ...) matches. Holding the source
also keeps its storage alive, so no later tensor can take its address while
the entry exists.
"""

from __future__ import annotations

from typing import Callable, Hashable

import torch


class Derived:
    """Named cache of tensors computed from a source tensor."""

    __slots__ = ("_entries",)

    def __init__(self) -> None:
        self._entries: dict[str, tuple[torch.Tensor, int, Hashable, object]] = {}

    def get(self, name: str, source: torch.Tensor, extra: Hashable, make: Callable[[], object]):
        """The entry `name` if it was made from `source` at its current version with `extra`,
        else `make()` (the stale entry is dropped first, so its memory can be reused)."""
        entry = self._entries.get(name)
        if entry is not None and entry[0] is source and entry[1] == source._version and entry[2] == extra:
            return entry[3]
        self._entries.pop(name, None)
        value = make()
        self._entries[name] = (source, source._version, extra, value)
        return value

    def drop(self, *names: str) -> None:
        """Forget the named entries (all of them when no name is given)."""
        if not names:
            self._entries.clear()
        for name in names:
            self._entries.pop(name, None)

    def drop_prefix(self, prefix: str) -> None:
        for name in [n for n in self._entries if n.startswith(prefix)]:
            del self._entries[name]

    def __contains__(self, name: str) -> bool:
        return name in self._entries
