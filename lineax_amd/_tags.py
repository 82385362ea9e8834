"""Operator tags (lineax/_tags.py:24-87).

Tags are unchecked declarations of structure; solvers and predicates read
them.  The same eight singletons as lineax, and the same transpose rule.
"""

from __future__ import annotations


class _Tag:
    __slots__ = ("name",)

    def __init__(self, name: str):
        self.name = name

    def __repr__(self) -> str:
        return self.name


symmetric_tag = _Tag("symmetric_tag")
diagonal_tag = _Tag("diagonal_tag")
tridiagonal_tag = _Tag("tridiagonal_tag")
unit_diagonal_tag = _Tag("unit_diagonal_tag")
lower_triangular_tag = _Tag("lower_triangular_tag")
upper_triangular_tag = _Tag("upper_triangular_tag")
positive_semidefinite_tag = _Tag("positive_semidefinite_tag")
negative_semidefinite_tag = _Tag("negative_semidefinite_tag")

# Tags that describe A and A^T alike (lineax/_tags.py:37-49).
_TRANSPOSE_INVARIANT = (
    symmetric_tag,
    unit_diagonal_tag,
    diagonal_tag,
    positive_semidefinite_tag,
    negative_semidefinite_tag,
    tridiagonal_tag,
)


def transpose_tags(tags: frozenset) -> frozenset:
    """Tags of the transpose of an operator carrying ``tags``.

    Symmetric operators keep every tag; otherwise transpose-invariant tags
    are kept and lower/upper triangularity swap (lineax/_tags.py:64-87).
    """
    if symmetric_tag in tags:
        return tags
    out = {t for t in _TRANSPOSE_INVARIANT if t in tags}
    if lower_triangular_tag in tags:
        out.add(upper_triangular_tag)
    if upper_triangular_tag in tags:
        out.add(lower_triangular_tag)
    return frozenset(out)


def as_tag_set(tags) -> frozenset:
    """A single tag or an iterable of tags, as a frozenset (lineax/_operator.py:61-67)."""
    if isinstance(tags, frozenset):
        return tags
    if isinstance(tags, _Tag):
        return frozenset([tags])
    try:
        return frozenset(iter(tags))
    except TypeError:
        return frozenset([tags])
