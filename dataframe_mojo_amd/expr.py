"""Expressions for the filter -> group_by -> agg path (expr.mojo:2869-3006).

Nodes: col, lit, comparisons (== != < <= > >=), %, & | ~, is_null,
is_not_null, is_in, is_between, arithmetic (+ - * /, unary -, abs), CASE
(when(...).then(...).otherwise(...), chained when/then), and the reductions
sum(min_count), count, len, min, max, mean, var(ddof), std(ddof) of a
column or of an arithmetic/CASE value, with alias.
"""
from __future__ import annotations

from dataclasses import dataclass

CMP_OPS = ("eq", "ne", "lt", "le", "gt", "ge")
AGG_OPS = ("len", "count", "sum", "min", "max", "mean", "var", "std")


@dataclass(frozen=True)
class Expr:
    kind: str
    args: tuple = ()
    name: str | None = None

    # naming ------------------------------------------------------------
    def alias(self, name: str) -> "Expr":
        return Expr(self.kind, self.args, name)

    def output_name(self) -> str:
        if self.name is not None:
            return self.name
        if self.kind == "col":
            return self.args[0]
        if self.kind == "agg" and self.args[0] == "len" and self.args[1] is None:
            return "len"
        for a in self.args:
            if isinstance(a, Expr):
                return a.output_name()
        return "literal"

    # comparisons ---------------------------------------------------------
    def _cmp(self, op: str, other) -> "Expr":
        return Expr("cmp", (op, self, _wrap(other)))

    def __eq__(self, other):  # type: ignore[override]
        return self._cmp("eq", other)

    def __ne__(self, other):  # type: ignore[override]
        return self._cmp("ne", other)

    def __lt__(self, other):
        return self._cmp("lt", other)

    def __le__(self, other):
        return self._cmp("le", other)

    def __gt__(self, other):
        return self._cmp("gt", other)

    def __ge__(self, other):
        return self._cmp("ge", other)

    def eq(self, other):
        return self._cmp("eq", other)

    def ne(self, other):
        return self._cmp("ne", other)

    __hash__ = object.__hash__

    # arithmetic / logic --------------------------------------------------
    def __mod__(self, other) -> "Expr":
        return Expr("mod", (self, _wrap(other)))

    def _arith(self, op: str, other, swap: bool = False) -> "Expr":
        a, b = (_wrap(other), self) if swap else (self, _wrap(other))
        return Expr("arith", (op, a, b))

    def __add__(self, other) -> "Expr":
        return self._arith("add", other)

    def __radd__(self, other) -> "Expr":
        return self._arith("add", other, True)

    def __sub__(self, other) -> "Expr":
        return self._arith("sub", other)

    def __rsub__(self, other) -> "Expr":
        return self._arith("sub", other, True)

    def __mul__(self, other) -> "Expr":
        return self._arith("mul", other)

    def __rmul__(self, other) -> "Expr":
        return self._arith("mul", other, True)

    def __truediv__(self, other) -> "Expr":
        """True division: Float64 (integers convert), IEEE by zero."""
        return self._arith("div", other)

    def __rtruediv__(self, other) -> "Expr":
        return self._arith("div", other, True)

    def __neg__(self) -> "Expr":
        return Expr("neg", (self,))

    def abs(self) -> "Expr":
        return Expr("abs", (self,))

    # CASE ----------------------------------------------------------------
    def otherwise(self, value) -> "Expr":
        """The ELSE of a when/then chain (without it, unmatched rows are
        null, expr.mojo:901-904)."""
        if self.kind != "case":
            raise ValueError("otherwise() follows when(...).then(...)")
        cond, then, other = self.args
        if other.kind == "case":
            return Expr("case", (cond, then, other.otherwise(value)), self.name)
        return Expr("case", (cond, then, _wrap(value)), self.name)

    def when(self, condition) -> "When":
        """A further WHEN of a chain: when(a).then(x).when(b).then(y)."""
        if self.kind != "case":
            raise ValueError("when() chains after when(...).then(...)")
        return When(_wrap(condition), self)

    def __and__(self, other) -> "Expr":
        return Expr("and", (self, _wrap(other)))

    def __or__(self, other) -> "Expr":
        return Expr("or", (self, _wrap(other)))

    def __invert__(self) -> "Expr":
        return Expr("not", (self,))

    def is_null(self) -> "Expr":
        return Expr("is_null", (self,))

    def is_not_null(self) -> "Expr":
        return Expr("is_not_null", (self,))

    def is_in(self, values) -> "Expr":
        """True when the value equals any candidate (integers and floats
        compare by value); a null candidate never matches, a null input stays
        null, and an empty list gives false (docs/expressions.md:445-448)."""
        return Expr("is_in", (self, tuple(values)))

    def is_between(self, lower, upper, closed: str = "both") -> "Expr":
        """lower <= value <= upper (closed "both", "left", "right" or
        "none" picks which bounds are inclusive), as the Kleene AND of the
        two comparisons (docs/expressions.md:448-450)."""
        if closed not in ("both", "left", "right", "none"):
            raise ValueError(f"closed must be both, left, right or none, not {closed!r}")
        return Expr("between", (self, _wrap(lower), _wrap(upper), closed))

    # reductions ----------------------------------------------------------
    def _agg(self, op: str, **kw) -> "Expr":
        return Expr("agg", (op, self, tuple(sorted(kw.items()))))

    def sum(self, min_count: int = 0) -> "Expr":
        """Skip nulls; zero when empty unless fewer than min_count are valid."""
        return self._agg("sum", min_count=min_count)

    def count(self) -> "Expr":
        """Number of non-null values, as Int64."""
        return self._agg("count")

    def len(self) -> "Expr":
        """Number of rows including nulls, as Int64."""
        return self._agg("len")

    def min(self) -> "Expr":
        """Smallest non-null value; NaN only when every valid value is NaN."""
        return self._agg("min")

    def max(self) -> "Expr":
        """Largest non-null value; NaN only when every valid value is NaN."""
        return self._agg("max")

    def mean(self) -> "Expr":
        """Arithmetic mean of non-null values as Float64; null when empty."""
        return self._agg("mean")

    def var(self, ddof: int = 1) -> "Expr":
        """Variance; null when fewer than ddof + 1 values."""
        return self._agg("var", ddof=ddof)

    def std(self, ddof: int = 1) -> "Expr":
        """Standard deviation; null when fewer than ddof + 1 values."""
        return self._agg("std", ddof=ddof)

    def __repr__(self) -> str:
        if self.kind == "col":
            s = f"col({self.args[0]!r})"
        elif self.kind == "lit":
            s = f"lit({self.args[0]!r})"
        elif self.kind == "cmp":
            s = f"({self.args[1]!r} {self.args[0]} {self.args[2]!r})"
        elif self.kind == "agg":
            s = f"{self.args[1]!r}.{self.args[0]}({dict(self.args[2])})" if self.args[1] is not None else "len()"
        else:
            s = f"{self.kind}({', '.join(repr(a) for a in self.args)})"
        return s if self.name is None else f"{s}.alias({self.name!r})"


class When:
    """when(condition) awaiting its then(value) (expr.mojo WHEN: left =
    predicate, right = then, extra = otherwise)."""

    def __init__(self, condition: Expr, chain: Expr | None = None):
        self.condition = condition
        self.chain = chain

    def then(self, value) -> Expr:
        case = Expr("case", (self.condition, _wrap(value), lit(None)))
        if self.chain is None:
            return case
        # a chained WHEN is the ELSE of the previous one
        return self.chain.otherwise(case)


def when(condition) -> When:
    return When(_wrap(condition))


def _wrap(v) -> Expr:
    return v if isinstance(v, Expr) else lit(v)


def col(name: str) -> Expr:
    return Expr("col", (name,))


def lit(value) -> Expr:
    return Expr("lit", (value,))


def len_() -> Expr:
    """len(): rows per group (the reference's `len()` expression)."""
    return Expr("agg", ("len", None, ()))


def sum_(name: str) -> Expr:
    return col(name).sum()
