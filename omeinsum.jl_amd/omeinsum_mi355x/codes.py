"""Contraction-code types and parsers: the host-side surface of OMEinsum kept unchanged.

Mirrors
  EinCode / DynamicEinCode / StaticEinCode   src/Core.jl:10-115
  ein"..." parser                            src/interfaces.jl:35-47
  get_size_dict! / get_size_dict             src/interfaces.jl:67-114
  parse_nested / filliys! / construct        src/einsequence.jl:9-153
  NestedEinsum (+ isleaf, siblings, ...)     src/einsequence.jl:161-244, :318-365
  SlicedEinsum                               src/slicing.jl:10-17, :139-153

Tensor indices of leaves are 0-based here (Julia's are 1-based).  Labels are any hashable
Python values (characters, ints, strings).  Error types follow the reference:
ArgumentError -> ``ArgumentError`` (a ValueError), DimensionMismatch -> ``DimensionMismatch``.
"""


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


def _unique(ix):
    out = []
    for l in ix:
        if l not in out:
            out.append(l)
    return out


class EinCode:
    """DynamicEinCode(ixs, iy) (src/Core.jl:82-110).  ``StaticEinCode`` is an alias: Python
    has no compile-time specialisation to gain from the static form."""

    def __init__(self, ixs, iy):
        ixs = [list(ix) for ix in ixs]
        if len(ixs) == 0:
            raise ValueError("number of input tensors must be greater than 0")
        self.ixs = ixs
        self.iy = list(iy)

    # -- AbstractEinsum interface
    def getixsv(self):
        return [list(ix) for ix in self.ixs]

    def getiyv(self):
        return list(self.iy)

    def labeltype(self):
        labels = [l for ix in self.ixs for l in ix] + list(self.iy)
        return type(labels[0]) if labels else object

    def uniquelabels(self):
        return _unique([l for ix in self.ixs for l in ix] + list(self.iy))

    def __eq__(self, other):
        return (isinstance(other, EinCode) and self.getixsv() == other.getixsv()
                and self.getiyv() == other.getiyv())

    def __hash__(self):
        return hash((tuple(tuple(ix) for ix in self.ixs), tuple(self.iy)))

    def replace(self, mapping):
        return EinCode([[mapping.get(l, l) for l in ix] for ix in self.ixs],
                       [mapping.get(l, l) for l in self.iy])

    def __repr__(self):
        j = lambda ix: "".join(map(str, ix)) if all(isinstance(l, str) and len(l) == 1 for l in ix) \
            else "∘".join(map(str, ix))
        return ", ".join(j(ix) for ix in self.ixs) + " -> " + j(self.iy)

    def __call__(self, *xs, size_info=None, **kw):
        from . import ein_ops as _e
        size_dict = get_size_dict(self.ixs, xs, size_info)
        return _e.einsum(self, xs, size_dict, **kw)


DynamicEinCode = EinCode
StaticEinCode = EinCode


def get_size_dict(ixs, xs, size_info=None):
    """get_size_dict!(ixs, xs, size_info) (src/interfaces.jl:67-106)."""
    d = dict(size_info) if size_info is not None else {}
    shapes = [tuple(_shape(x)) for x in xs]
    if len(shapes) < 1:
        raise ValueError("empty input tensors")
    if len(ixs) != len(shapes):
        raise ArgumentError("%d tensors labelled by %d indices" % (len(shapes), len(ixs)))
    for ix, s in zip(ixs, shapes):
        if len(ix) != len(s):
            raise ArgumentError("indices %r invalid for tensor with ndims = %d" % (tuple(ix), len(s)))
        for k, n in zip(ix, s):
            if k in d:
                if d[k] != n:
                    raise DimensionMismatch("%r = %d or %d)?" % (k, d[k], n))
            else:
                d[k] = int(n)
    return d


def _shape(x):
    s = getattr(x, "shape", None)
    if s is None:
        return ()
    return tuple(int(v) for v in s)


# ---------------------------------------------------------------------------------------
# string front end (src/interfaces.jl:35-47) and nested parser (src/einsequence.jl:9-153)


def _is_label(c):
    return c.isalpha()


def ein(s):
    """ein"..." (src/interfaces.jl:35-47).  Returns an EinCode, or a NestedEinsum when the
    input part has parentheses."""
    s = s.replace("\n", "").replace(" ", "")
    if s.count("->") != 1:
        raise ArgumentError("invalid einsum specification %s" % s)
    sixs, siy = s.split("->")
    if not all(_is_label(c) or c in "(),"for c in sixs) or not all(_is_label(c) for c in siy):
        raise ArgumentError("invalid einsum specification %s" % s)
    if "(" in sixs:
        return parse_nested(sixs, list(siy))
    iy = list(siy)
    ixs = [list(ix) for ix in sixs.split(",")]
    return EinCode(ixs, iy)


class _IndexGroup:
    def __init__(self, n):
        self.inds = []
        self.n = n


class _Constructor:
    def __init__(self):
        self.args = []
        self.inds = []
        self.iy = []


def _union_into(dst, src):
    for l in src:
        if l not in dst:
            dst.append(l)


def parse_nested(s, iy=()):
    """parse_nested (src/einsequence.jl:9-25), parse_parens (:34-72), filliys! (:80-94),
    construct (:254-260)."""
    if s.count("(") != s.count(")"):
        raise ArgumentError("Parentheses don't pair up in %s" % s)

    def parse_parens(i, narg):
        out = _Constructor()
        g = _IndexGroup(narg)
        while i < len(s):
            c = s[i]
            j = i + 1
            if c == "(":
                j, out2, narg = parse_parens(j, narg)
                out.args.append(out2)
                _union_into(out.inds, out2.inds)
            elif c in "),":
                if g.inds:
                    out.args.append(g)
                    _union_into(out.inds, g.inds)
                if c == ",":
                    narg += 1
                    g = _IndexGroup(narg)
                else:
                    return j, out, narg
            elif _is_label(c):
                g.inds.append(c)
            else:
                raise ArgumentError("parsing %s failed, %s is not a valid entry" % (s, c))
            i = j
        if g.inds:
            out.args.append(g)
            _union_into(out.inds, g.inds)
        return i, out, narg

    res = parse_parens(0, 0)
    out = res[1]
    out.iy.extend(iy)
    _filliys(out)
    return _construct(out)


def _filliys(ne):
    for i, arg in enumerate(ne.args):
        if isinstance(arg, _IndexGroup):
            continue
        _union_into(arg.iy, [l for l in arg.inds if l in ne.iy])
        for j, other in enumerate(ne.args):
            if i == j:
                continue
            _union_into(arg.iy, [l for l in arg.inds if l in other.inds])
        _filliys(arg)
    return ne


def _construct(ne):
    ixs = [list(a.inds) if isinstance(a, _IndexGroup) else list(a.iy) for a in ne.args]
    eins = EinCode(ixs, list(ne.iy))
    args = [_construct(a) if isinstance(a, _Constructor) else NestedEinsum.leaf(a.n)
            for a in ne.args]
    return NestedEinsum(args, eins)


# ---------------------------------------------------------------------------------------
# contraction trees


class NestedEinsum:
    """DynamicNestedEinsum (src/einsequence.jl:161-185): a leaf (tensorindex >= 0) or an
    internal node with ``args`` and an ``eins`` EinCode."""

    def __init__(self, args, eins):
        args = list(args)
        if len(args) != len(eins.ixs):
            raise ValueError("number of siblings does not match the code")
        self.args = args
        self.tensorindex = -1
        self.eins = eins

    @classmethod
    def leaf(cls, i):
        obj = cls.__new__(cls)
        obj.args = []
        obj.tensorindex = int(i)
        obj.eins = None
        return obj

    def isleaf(self):
        return self.tensorindex != -1

    def siblings(self):
        return self.args

    def rootcode(self):
        return self.eins

    def __eq__(self, other):
        if not isinstance(other, NestedEinsum) or self.isleaf() != other.isleaf():
            return False
        if self.isleaf():
            return self.tensorindex == other.tensorindex
        return (self.eins == other.eins and len(self.args) == len(other.args)
                and all(a == b for a, b in zip(self.args, other.args)))

    def __hash__(self):
        # a tree is not modified after construction: its hash is computed once
        h = self.__dict__.get("_hash")
        if h is None:
            h = hash(self.to_plain_key())
            self.__dict__["_hash"] = h
        return h

    def to_plain_key(self):
        if self.isleaf():
            return self.tensorindex
        return (tuple(a.to_plain_key() for a in self.args),
                tuple(tuple(ix) for ix in self.eins.ixs), tuple(self.eins.iy))

    def to_plain(self):
        """Plain form (int leaf | (args, ixs, iy)) used by the tree wire format."""
        if self.isleaf():
            return self.tensorindex
        return ([a.to_plain() for a in self.args], [list(ix) for ix in self.eins.ixs],
                list(self.eins.iy))

    @classmethod
    def from_plain(cls, p):
        if isinstance(p, int):
            return cls.leaf(p)
        args, ixs, iy = p
        return cls([cls.from_plain(a) for a in args], EinCode(ixs, iy))

    # -- AbstractEinsum interface (src/einsequence.jl:318-365)
    def collect_ixs(self, d=None):
        d = {} if d is None else d
        for a, ix in zip(self.args, self.eins.ixs):
            if a.isleaf():
                d[a.tensorindex] = list(ix)
            else:
                a.collect_ixs(d)
        return d

    def getixsv(self):
        if self.isleaf():
            raise RuntimeError("Can not call `getixsv` on leaf nodes!")
        d = self.collect_ixs()
        return [d[i] for i in sorted(d)]

    def getiyv(self):
        if self.isleaf():
            raise RuntimeError("Can not call `getiyv` on leaf nodes!")
        return list(self.eins.iy)

    def uniquelabels(self):
        ul = self.__dict__.get("_labels")
        if ul is None:
            ul = _unique([l for ix in self.getixsv() for l in ix] + self.getiyv())
            self.__dict__["_labels"] = ul
        return list(ul)

    def labeltype(self):
        return EinCode(self.getixsv(), self.getiyv()).labeltype()

    def flatten(self):
        """flatten (src/einsequence.jl:326-343)."""
        return EinCode(self.getixsv(), self.getiyv())

    def replace(self, mapping):
        if self.isleaf():
            return NestedEinsum.leaf(self.tensorindex)
        return NestedEinsum([a.replace(mapping) for a in self.args], self.eins.replace(mapping))

    def internal_nodes(self):
        """Post-order list of internal nodes (the evaluation order of
        src/einsequence.jl:279-287)."""
        out = []

        def rec(n):
            if n.isleaf():
                return
            for a in n.args:
                rec(a)
            out.append(n)
        rec(self)
        return out

    def __repr__(self):
        if self.isleaf():
            return "Leaf(%d)" % self.tensorindex
        return "(" + ", ".join(repr(a) for a in self.args) + " | " + repr(self.eins) + ")"

    def __call__(self, *xs, size_info=None, **kw):
        from . import ein_ops as _e
        size_dict = get_size_dict(self.getixsv(), xs, size_info)
        return _e.einsum(self, xs, size_dict, **kw)


DynamicNestedEinsum = NestedEinsum
StaticNestedEinsum = NestedEinsum


class SlicedEinsum:
    """SlicedEinsum(slicing, eins) (src/slicing.jl:10-17)."""

    def __init__(self, slicing, eins):
        self.slicing = list(slicing)
        self.eins = eins

    def __eq__(self, other):
        return (isinstance(other, SlicedEinsum) and self.slicing == other.slicing
                and self.eins == other.eins)

    def __hash__(self):
        return hash((tuple(self.slicing), self.eins))

    def getixsv(self):
        return self.eins.getixsv()

    def getiyv(self):
        return self.eins.getiyv()

    def uniquelabels(self):
        return self.eins.uniquelabels()

    def labeltype(self):
        return self.eins.labeltype()

    def flatten(self):
        return self.eins.flatten()

    def replace(self, mapping):
        return SlicedEinsum([mapping.get(l, l) for l in self.slicing], self.eins.replace(mapping))

    def __repr__(self):
        return "SlicedEinsum(%r, %r)" % (self.slicing, self.eins)

    def __call__(self, *xs, size_info=None, **kw):
        from . import ein_ops as _e
        size_dict = get_size_dict(self.getixsv(), xs, size_info)
        return _e.einsum(self, xs, size_dict, **kw)


def drop_slicedim(ne, slices):
    """drop_slicedim (src/slicing.jl:139-145)."""
    if ne.isleaf():
        return ne
    ixs = [[l for l in ix if l not in slices] for ix in ne.eins.ixs]
    iy = [l for l in ne.eins.iy if l not in slices]
    return NestedEinsum([drop_slicedim(a, slices) for a in ne.args], EinCode(ixs, iy))


def getixsv(code):
    return code.getixsv()


def getiyv(code):
    return code.getiyv()


def uniquelabels(code):
    return code.uniquelabels()


def labeltype(code):
    return code.labeltype()


def flatten(code):
    return code if isinstance(code, EinCode) else code.flatten()


def uniformsize(code, size):
    """uniformsize (src/contractionorder.jl:36)."""
    return {l: size for l in code.uniquelabels()}
