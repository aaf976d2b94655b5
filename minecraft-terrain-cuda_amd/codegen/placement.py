"""Placement of density-graph values into the nested control scopes of one generated function.

A generated evaluator is one straight-line function. Every value it needs is a key
(node, coordinate variant, semantics) bound to an SSA local at its first demand, in the scope
that demands it. Lazy nodes (lerp, range choice, interval select, the point path's zero and
range short-circuits) open child scopes, one per arm of a branch; a loop (the find_top_surface
probe) opens a repeating scope. A value already bound in the demanding scope or an enclosing
one is reused; otherwise it is computed again where it is demanded.

Computing a value again is free when the two places are arms of one branch (only one runs),
so the builder keeps that. Two cases would compute a value twice on one path: a second demand
in a scope that is not an arm sibling of the first (e.g. first inside a branch, then after
it), and a loop-invariant value demanded inside a repeating scope. Those keys are memoised:
a closure with a computed-flag, defined at the top of the function (or of the loop body, if
the value depends on the loop variable), computing the value at its first call. The decision
is a fixed point: build, collect the keys that need a memo, rebuild with them, until no new
key appears (memoising a key moves its operands' demands into its closure).

Every scope carries the count of operations it performs (by class: FP64, FP32, INT32, memory
reads), emitted as an `MCT_OPS` hook at its top: the host build that counts operations adds
them each time the scope runs, the device build ignores it.
"""
from collections import defaultdict

CLASSES = ('f64', 'f32', 'i32', 'mem')


def ops(f64=0, f32=0, i32=0, mem=0):
    return (f64, f32, i32, mem)


def add_ops(a, b, k=1):
    return tuple(x + k * y for x, y in zip(a, b))


class Branch:
    """Identity of one branch construct; its arms are mutually exclusive scopes."""
    __slots__ = ()


class Scope:
    def __init__(self, parent, branch=None, loop_var=None, outer=None):
        self.parent = parent
        self.branch = branch
        self.loop_var = loop_var
        self.outer = outer         # for a closure body: the scope whose prologue defines it
        self.items = []            # str lines, or ('branch', [(header, Scope)]), ('loop', header, Scope)
        self.prologue = []         # memo closures defined at the top of this scope
        self.memo_names = {}
        self.avail = {}
        self.ops = ops()
        self.root = parent.root if parent is not None else self
        self.chain = (parent.chain if parent is not None else ()) + (self,)

    def enclosing(self):
        """Scopes this one runs inside, innermost first: parents, and past a closure body the
        scope that defines the closure."""
        s = self
        while s is not None:
            yield s
            s = s.parent if s.parent is not None else s.outer

    def add(self, line, cost=None):
        self.items.append(line)
        if cost:
            self.ops = add_ops(self.ops, cost)


def exclusive(a, b):
    """True when scopes a and b can never both run in one call: they lie in different arms
    of one branch construct."""
    if a.root is not b.root:
        return False
    n = min(len(a.chain), len(b.chain))
    i = 0
    while i < n and a.chain[i] is b.chain[i]:
        i += 1
    if i == len(a.chain) or i == len(b.chain):
        return False
    ca, cb = a.chain[i], b.chain[i]
    return ca.branch is not None and ca.branch is cb.branch


class FunctionBuilder:
    """Builds one function body. `lower` supplies alias(key) (the key a pass-through node
    forwards to, or None), literal(key), emit(builder, key, scope) -> bound name, and
    invariant(key, loop_var)."""

    def __init__(self, lower, memo):
        self.lower = lower
        self.memo = memo
        self.want = set()
        self.emitted = defaultdict(list)
        self.n = 0
        self.root = Scope(None)

    def fresh(self, prefix='v'):
        self.n += 1
        return '%s%d' % (prefix, self.n)

    def value(self, key, scope):
        while True:
            to = self.lower.alias(key)
            if to is None:
                break
            key = to
        lit = self.lower.literal(key)
        if lit is not None:
            return lit
        s = scope
        while s is not None:
            if key in s.avail:
                return s.avail[key]
            s = s.parent
        if key in self.memo:
            return self._memo_call(key, scope)
        for prev in self.emitted[key]:
            if not exclusive(prev, scope):
                self.want.add(key)
        for s in scope.enclosing():
            if s.loop_var is not None and self.lower.invariant(key, s.loop_var):
                self.want.add(key)
        self.emitted[key].append(scope)
        name = self.lower.emit(self, key, scope)
        scope.avail[key] = name
        return name

    def bind(self, scope, expr, cost=None, ctype='const float'):
        name = self.fresh()
        scope.add('%s %s = %s;' % (ctype, name, expr), cost)
        return name

    def branch(self, scope, result, arms, cost=None):
        """arms: [(header, fill)] where fill(arm_scope) returns the arm's value expression;
        the branch assigns it to `result` (declared by the caller)."""
        b = Branch()
        built = []
        for header, fill in arms:
            s = Scope(scope, branch=b)
            v = fill(s)
            s.add('%s = %s;' % (result, v))
            built.append((header, s))
        scope.items.append(('branch', built))
        if cost:
            scope.ops = add_ops(scope.ops, cost)

    def loop(self, scope, header, var, fill):
        s = Scope(scope, loop_var=var)
        fill(s)
        scope.items.append(('loop', header, s))

    def _home(self, key, scope):
        last = scope
        for s in scope.enclosing():
            if s.loop_var is not None and not self.lower.invariant(key, s.loop_var):
                return s
            last = s
        return last

    def _memo_call(self, key, scope):
        home = self._home(key, scope)
        name = home.memo_names.get(key)
        if name is None:
            name = self.fresh('g')
            home.memo_names[key] = name
            body = Scope(None, outer=home)
            v = self.lower.emit(self, key, body)
            home.prologue.append((name, body, v))
        return name + '()'


def render_scope(s, indent):
    pad = '    ' * indent
    out = []
    if any(s.ops):
        out.append(pad + 'MCT_OPS(%d, %d, %d, %d);' % s.ops)
    for name, body, v in s.prologue:
        out.append(pad + 'float m_%s = 0.0f; bool h_%s = false;' % (name, name))
        out.append(pad + 'auto %s = [&]() MCT_CLOSURE -> float {' % name)
        out.append(pad + '    if (!h_%s) {' % name)
        out += render_scope(body, indent + 2)
        out.append(pad + '        m_%s = %s;' % (name, v))
        out.append(pad + '        h_%s = true;' % name)
        out.append(pad + '    }')
        out.append(pad + '    return m_%s;' % name)
        out.append(pad + '};')
    for it in s.items:
        if isinstance(it, str):
            out.append(pad + it)
        elif it[0] == 'branch':
            for k, (header, arm) in enumerate(it[1]):
                out.append(pad + ('%s {' % header if k == 0 else '} %s {' % header))
                out += render_scope(arm, indent + 1)
            out.append(pad + '}')
        else:
            _, header, body = it
            out.append(pad + '%s {' % header)
            out += render_scope(body, indent + 1)
            out.append(pad + '}')
    return out


def scope_ops(s, all_paths):
    """Operations of a scope: its own plus its closures' (run once), plus, with all_paths,
    every arm and loop body (once each)."""
    total = s.ops
    for _, body, _ in s.prologue:
        total = add_ops(total, scope_ops(body, all_paths))
    if all_paths:
        for it in s.items:
            if isinstance(it, str):
                continue
            if it[0] == 'branch':
                for _, arm in it[1]:
                    total = add_ops(total, scope_ops(arm, True))
            else:
                total = add_ops(total, scope_ops(it[2], True))
    return total


def build(lower, fill):
    """Runs fill(builder, root_scope) to the memo fixed point; returns the final builder."""
    memo = set()
    while True:
        b = FunctionBuilder(lower, frozenset(memo))
        fill(b, b.root)
        if b.want <= memo:
            return b
        memo |= b.want
