"""A subset of the Bmad lattice language to the element table (bmad/parsing/bmad_parser.f90:33).

Reads ``parameter[...]`` (p0c or e_tot, particle), variables, element definitions (by type or
by inheriting another element's attributes), ``name[attr] = value`` settings, lines with
repetition (``3*cell``), reflection (``-cell``) and nesting, ``use`` and ``call, file = ...``, with arithmetic
expressions over numbers, constants, variables and ``ele[attr]`` references. Element types are
those of this path (drift, quadrupole, sbend, rbend, sextupole, octupole, marker, rfcavity,
kickers, monitor, instrument, collimators, pipe, thick_multipole), matched by unique prefix as
the reference parser does. An rbend becomes an sbend as the reference's bookkeeping makes it
(arc length from the chord, edge angles increased by half the bend angle). Overlays set the
attributes they control to the sum of their expressions at the overlay variables' values
(slave names may use the wildcards % and *); groups, which only change attributes when their
variables change after the lattice is set up, change nothing at load. Superposition, reversed
elements (``--``) and attributes outside this path raise ``LatticeFileError``.
The result goes through ``elements.make_element`` (Bmad defaults and step bookkeeping).
"""
from __future__ import annotations

import ast
import math
import operator
import os
import re

from . import elements as E


class LatticeFileError(ValueError):
    pass


_TYPES = {
    "drift": E.DRIFT, "quadrupole": E.QUADRUPOLE, "sbend": E.SBEND, "rbend": E.SBEND,
    "sextupole": E.SEXTUPOLE, "octupole": E.OCTUPOLE, "marker": E.MARKER, "rfcavity": E.RFCAVITY,
    "hkicker": E.HKICKER, "vkicker": E.VKICKER, "kicker": E.KICKER, "monitor": E.MONITOR,
    "instrument": E.INSTRUMENT, "ecollimator": E.ECOLLIMATOR, "rcollimator": E.RCOLLIMATOR,
    "pipe": E.PIPE, "thick_multipole": E.THICK_MULTIPOLE, "multipole": E.MULTIPOLE,
    "ab_multipole": E.AB_MULTIPOLE,
}

# particle_species_mod.f90 masses [eV] and charges
_PARTICLES = {"electron": (0.51099895069e6, -1), "positron": (0.51099895069e6, 1),
              "proton": (938.27208816e6, 1), "antiproton": (938.27208816e6, -1),
              "muon": (105.6583755e6, -1), "antimuon": (105.6583755e6, 1)}

_CONSTANTS = {"pi": math.pi, "twopi": 2 * math.pi, "fourpi": 4 * math.pi, "e_log": math.e,
              "sqrt_2": math.sqrt(2), "degrees": 180 / math.pi, "raddeg": 180 / math.pi,
              "c_light": E.C_LIGHT, "m_electron": E.M_ELECTRON, "m_proton": 938.27208816e6,
              "e_charge": 1.602176634e-19, "true": 1.0, "false": 0.0}

_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin,
          "acos": math.acos, "atan": math.atan, "atan2": math.atan2, "sqrt": math.sqrt,
          "exp": math.exp, "log": math.log, "abs": abs, "sinh": math.sinh, "cosh": math.cosh,
          "tanh": math.tanh}

# lattice attribute -> make_element keyword (same name unless listed)
_ATTRS = {"l", "k1", "k2", "k3", "g", "dg", "e1", "e2", "fint", "fintx", "hgap", "hgapx", "hkick",
          "vkick", "kick", "tilt", "ref_tilt", "x_offset", "y_offset", "z_offset", "x_pitch",
          "y_pitch", "voltage", "rf_frequency", "phi0", "phi0_multipass", "x_limit", "y_limit",
          "x1_limit", "x2_limit", "y1_limit", "y2_limit", "num_steps", "ds_step", "angle", "rho",
          "aperture_type", "aperture_at", "fringe_type", "fringe_at", "harmon", "r0_mag",
          "r0_elec", "fq1", "fq2", "exact_multipoles"}
# multipoles: a0..a21, b0..b21 (magnetic) and a0_elec..b21_elec (electric)
_MULTIPOLE = re.compile(r"^([ab])(\d+)(_elec)?$")
# a thin multipole's integrated strengths k0l..k21l, skew strengths k0sl..k21sl and order tilts
# t0..t21
_THIN = re.compile(r"^(?:k(\d+)(s?)l|t(\d+))$")
_ENUMS = {"aperture_type": {"rectangular": E.APERTURE_RECTANGULAR, "elliptical": E.APERTURE_ELLIPTICAL},
          "aperture_at": {"entrance_end": E.ENTRANCE_END, "exit_end": E.EXIT_END,
                          "both_ends": E.BOTH_ENDS, "no_end": E.NO_END},
          "fringe_type": {"none": E.FRINGE_NONE, "hard_edge_only": E.FRINGE_HARD_EDGE_ONLY,
                          "basic_bend": E.FRINGE_BASIC_BEND, "full": E.FRINGE_FULL,
                          "soft_edge_only": E.FRINGE_SOFT_EDGE_ONLY},
          "fringe_at": {"entrance_end": E.ENTRANCE_END, "exit_end": E.EXIT_END,
                        "both_ends": E.BOTH_ENDS, "no_end": E.NO_END},
          "exact_multipoles": {"off": E.EXACT_OFF, "horizontally_pure": E.EXACT_HORIZONTALLY_PURE,
                               "vertically_pure": E.EXACT_VERTICALLY_PURE}}
_IGNORED = ("beginning", "bmad_com", "particle_start", "space_charge_com", "no_digested",
            "title", "expand_lattice", "end_file", "debug_marker")


def _lookup_type(word: str):
    w = word.lower()
    if w in _TYPES:
        return w
    hits = [t for t in _TYPES if t.startswith(w)]
    return hits[0] if len(hits) == 1 else None


class _Parser:
    def __init__(self):
        self.params = {}
        self.vars = dict(_CONSTANTS)
        self.eles = {}    # name -> (type name, {attr: value})
        self.lines = {}   # name -> list of items
        self.use = None
        self.start_at = None  # start_branch_at: the element the branch starts with
        # name -> {"kind": "overlay" | "group", "slaves": [(pattern, attr, expr)], "vars": {}}
        self.controllers = {}

    # -- expressions ----------------------------------------------------------------------
    def value(self, text: str):
        t = text.strip()
        t = re.sub(r"([A-Za-z_][\w.]*)\[([A-Za-z_]\w*)\]", lambda m: f"__ref__('{m.group(1)}','{m.group(2)}')", t)
        t = t.replace("^", "**")
        t = re.sub(r"\.(true|false)\.", lambda m: "1" if m.group(1).lower() == "true" else "0", t, flags=re.I)
        try:
            tree = ast.parse(t, mode="eval")
        except SyntaxError as e:
            raise LatticeFileError(f"bad expression {text!r}") from e
        return self._eval(tree.body, text)

    def _eval(self, n, text):
        ops = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
               ast.Div: operator.truediv, ast.Pow: operator.pow}
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.BinOp) and type(n.op) in ops:
            return ops[type(n.op)](self._eval(n.left, text), self._eval(n.right, text))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = self._eval(n.operand, text)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.Name):
            k = n.id.lower()
            if k in self.vars:
                return self.vars[k]
            raise LatticeFileError(f"undefined name {n.id!r} in {text!r}")
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name):
            f = n.func.id
            if f == "__ref__":
                ele, attr = (a.value.lower() for a in n.args)
                return self._attr(ele, attr)
            if f.lower() in _FUNCS:
                return _FUNCS[f.lower()](*[self._eval(a, text) for a in n.args])
        raise LatticeFileError(f"unsupported expression {text!r}")

    def _attr(self, ele, attr):
        if ele == "parameter":
            return self.params[attr]
        if ele not in self.eles:
            raise LatticeFileError(f"unknown element {ele!r}")
        return self.eles[ele][1].get(attr, 0.0)

    # -- statements -----------------------------------------------------------------------
    def statements(self, path):
        text = open(path).read()
        out, cur, depth = [], "", 0
        for raw in text.splitlines():
            line = raw.split("!", 1)[0].rstrip()
            if not line.strip():
                continue
            cur = (cur + " " + line.strip()) if cur else line.strip()
            depth = cur.count("(") + cur.count("{") - cur.count(")") - cur.count("}")
            if cur.endswith("&"):
                cur = cur[:-1]
                continue
            if cur.endswith(",") or depth > 0:
                continue
            out.append(cur)
            cur = ""
        if cur:
            out.append(cur)
        return out

    def parse(self, path):
        base = os.path.dirname(os.path.abspath(path))
        for st in self.statements(path):
            low = st.lower()
            if low.startswith("call"):
                m = re.match(r"call\s*,\s*file\s*=\s*(\S+)", st, re.I)
                if not m:
                    raise LatticeFileError(f"bad call statement {st!r}")
                self.parse(os.path.join(base, m.group(1).strip("'\"")))
                continue
            if low.startswith("use"):
                self.use = re.split(r"\s*,\s*", st, maxsplit=1)[1].strip().lower()
                continue
            m = re.match(r"start_branch_at\s*,?\s*([\w.]+)\s*$", st, re.I)
            if m:
                self.start_at = m.group(1).lower()
                continue
            m = re.match(r"beginning\s*\[\s*(e_tot|p0c)\s*\]\s*=\s*(.+)$", st, re.I)
            if m:
                # the reference energy of an open lattice, as parameter[e_tot] / [p0c]
                self.params.setdefault(m.group(1).lower(), self.value(m.group(2)))
                continue
            if any(low.startswith(k) for k in _IGNORED):
                continue
            m = re.match(r"parameter\s*\[\s*(\w+)\s*\]\s*=\s*(.+)$", st, re.I)
            if m:
                k, v = m.group(1).lower(), m.group(2).strip()
                self.params[k] = v.strip("'\"").lower() if k in ("particle", "geometry") else self.value(v)
                continue
            m = re.match(r"(\w+)::([\w.%*]+)\s*\[\s*(\w+)\s*\]\s*=\s*(.+)$", st)
            if m:
                # type::name-pattern[attr] = value: every element of that type whose name
                # matches (wildcards % and *); %[attr] in the value is the element's own value
                tname = _lookup_type(m.group(1))
                if tname is None:
                    raise LatticeFileError(f"unknown element type in {st!r}")
                rx = re.compile("^" + re.escape(m.group(2).lower()).replace("%", ".").replace("\\*", ".*") + "$")
                self._set_class(tname, rx, m.group(3).lower(), m.group(4), st)
                continue
            m = re.match(r"([\w.]+)\s*\[\s*(\w+)\s*\]\s*=\s*(.+)$", st)
            if m and m.group(1).lower() not in self.eles and m.group(1).lower() not in self.controllers \
                    and _lookup_type(m.group(1)) is not None:
                # type[attr] = value: every element of that type defined so far
                self._set_class(_lookup_type(m.group(1)), None, m.group(2).lower(), m.group(3), st)
                continue
            if m:
                name, attr = m.group(1).lower(), m.group(2).lower()
                if name in self.controllers:
                    ctl = self.controllers[name]
                    if attr not in ctl["vars"]:
                        raise LatticeFileError(f"{name} has no variable {attr!r}: {st!r}")
                    ctl["vars"][attr] = self.value(m.group(3))
                    continue
                if name not in self.eles:
                    raise LatticeFileError(f"attribute of unknown element in {st!r}")
                self._set(self.eles[name][1], attr, m.group(3))
                continue
            m = re.match(r"([\w.]+)\s*:\s*(.*)$", st)
            if m:
                self._define(m.group(1).lower(), m.group(2), st)
                continue
            m = re.match(r"([A-Za-z_]\w*)\s*=\s*(.+)$", st)
            if m:
                self.vars[m.group(1).lower()] = self.value(m.group(2))
                continue
            raise LatticeFileError(f"statement not understood: {st!r}")

    def _define(self, name, rest, st):
        parts = _split_top(rest)
        head = parts[0].strip()
        m = re.match(r"line\s*=\s*\((.*)\)\s*$", rest, re.I | re.S)
        if head.lower().startswith("line") and m:
            self.lines[name] = _split_top(m.group(1))
            return
        mc = re.match(r"(overlay|group)\s*=\s*\{(.*)\}\s*$", head, re.I | re.S)
        if mc:
            self._controller(name, mc.group(1).lower(), mc.group(2), parts[1:], st)
            return
        if head.lower() in ("list", "overlay", "group", "girder") or "superimpose" in rest.lower():
            raise LatticeFileError(f"{head} elements are not supported: {st!r}")
        # an element defined earlier takes precedence over a type name it abbreviates
        # ("rfcav" after "rfcav: rfcavity, ...")
        if head.lower() in self.eles:
            tname, parent = self.eles[head.lower()]
            attrs = dict(parent)
        elif _lookup_type(head) is not None:
            tname = _lookup_type(head)
            attrs = {}
        else:
            raise LatticeFileError(f"element type {head!r} is not on this path: {st!r}")
        for p in parts[1:]:
            if "=" in p:
                k, v = p.split("=", 1)
                self._set(attrs, k.strip().lower(), v)
            elif p.strip().lower() == "tilt":
                attrs["__default_tilt__"] = True
            else:
                raise LatticeFileError(f"attribute {p.strip()!r} not supported: {st!r}")
        self.eles[name] = (tname, attrs)

    def _controller(self, name, kind, body, rest, st):
        """overlay / group: slaves 'name[attr]: expression' (an omitted expression is the first
        variable), 'var = {...}' and initial variable values."""
        variables = {}
        others = []
        for p in rest:
            m = re.match(r"var\s*=\s*\{(.*)\}$", p.strip(), re.I | re.S)
            if m:
                for v in _split_top(m.group(1)):
                    if v:
                        variables[v.lower()] = 0.0
            elif p.strip():
                others.append(p)
        if not variables:
            raise LatticeFileError(f"{kind} without var = {{...}}: {st!r}")
        for p in others:
            k, _, v = p.partition("=")
            k = k.strip().lower()
            if k == "type":
                continue
            if k not in variables or not v:
                raise LatticeFileError(f"{kind} attribute {p.strip()!r} not supported: {st!r}")
            variables[k] = self.value(v)
        slaves = []
        for item in _split_top(body):
            if not item:
                continue
            m = re.match(r"([\w.%*]+)\s*\[\s*(\w+)\s*\]\s*(?::\s*(.+))?$", item.strip(), re.S)
            if not m:
                raise LatticeFileError(f"{kind} slave {item!r} not supported: {st!r}")
            attr = m.group(2).lower()
            if attr not in _ATTRS and not _MULTIPOLE.match(attr) and not _THIN.match(attr):
                raise LatticeFileError(f"{kind} controls attribute {attr!r}, not on this path")
            expr = m.group(3) if m.group(3) is not None else next(iter(variables))
            slaves.append((m.group(1).lower(), attr, expr))
        self.controllers[name] = {"kind": kind, "slaves": slaves, "vars": variables}

    def apply_overlays(self):
        """Each attribute an overlay controls becomes the sum of its overlays' expressions at
        their variables' values (bookkeeper_mod.f90:1779-1831, overlay_change_this); groups
        change attributes only when their variables change later, so they change nothing
        here."""
        sums = {}
        for ctl in self.controllers.values():
            if ctl["kind"] != "overlay":
                continue
            saved = dict(self.vars)
            self.vars.update(ctl["vars"])
            try:
                for pattern, attr, expr in ctl["slaves"]:
                    rx = re.compile("^" + re.escape(pattern).replace("%", ".").replace("\\*", ".*") + "$")
                    hits = [n for n in self.eles if rx.match(n)]
                    if not hits:
                        raise LatticeFileError(f"overlay slave {pattern!r} matches no element")
                    val = self.value(expr)
                    for n in hits:
                        sums[(n, attr)] = sums.get((n, attr), 0.0) + val
            finally:
                self.vars = saved
        for (n, attr), val in sums.items():
            self.eles[n][1][attr] = val

    def _set_class(self, tname, rx, attr, text, st):
        hits = [n for n, (t, _) in self.eles.items() if t == tname and (rx is None or rx.match(n))]
        if rx is not None and not hits:
            raise LatticeFileError(f"no element matches {st!r}")
        for n in hits:
            self._set(self.eles[n][1], attr, text.replace("%[", f"{n}["))

    def _set(self, attrs, k, v):
        if _MULTIPOLE.match(k) or _THIN.match(k):
            attrs[k] = self.value(v)
            return
        if k not in _ATTRS:
            raise LatticeFileError(f"attribute {k!r} is not supported")
        if k in _ENUMS:
            key = v.strip().strip("'\"").lower()
            if key not in _ENUMS[k]:
                raise LatticeFileError(f"{k} = {v.strip()} is not supported")
            attrs[k] = _ENUMS[k][key]
        else:
            attrs[k] = self.value(v)

    # -- expansion ------------------------------------------------------------------------
    def expand(self, name, depth=0):
        if depth > 50:
            raise LatticeFileError("line nesting too deep")
        if name in self.eles:
            return [name]
        if name not in self.lines:
            raise LatticeFileError(f"unknown line or element {name!r}")
        out = []
        for item in self.lines[name]:
            it = item.strip().lower()
            if not it:
                continue
            m = re.match(r"(\d+)\s*\*\s*(.+)$", it)
            n, body = (int(m.group(1)), m.group(2).strip()) if m else (1, it)
            if body.startswith("--"):
                raise LatticeFileError(f"reversed elements are not supported: {item!r}")
            reflect = body.startswith("-")
            if reflect:
                body = body[1:].strip()
            if body.startswith("("):
                sub = f"__anon{len(self.lines)}"
                self.lines[sub] = _split_top(body[1:-1])
                body = sub
            seq = self.expand(body, depth + 1)
            # reflection (a leading '-'): the same elements in reverse order
            out += (seq[::-1] if reflect else seq) * n
        return out


def _split_top(s: str):
    """Split at commas outside parentheses and braces."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return [x.strip() for x in out]


def load(path: str):
    """Parse a lattice file; returns (element records (make_element), mc2 [eV], charge)."""
    P = _Parser()
    P.parse(path)
    species = str(P.params.get("particle", "electron")).lower()
    if species not in _PARTICLES:
        raise LatticeFileError(f"particle {species!r} is not supported")
    mc2, charge = _PARTICLES[species]
    if "p0c" in P.params:
        p0c = P.params["p0c"]
        e_tot = E.p0c_to_e_tot(p0c, mc2)
    elif "e_tot" in P.params:
        e_tot = P.params["e_tot"]
        p0c = E.e_tot_to_p0c(e_tot, mc2)
    else:
        raise LatticeFileError("parameter[p0c] or parameter[e_tot] is required")
    use = P.use or (list(P.lines)[-1] if P.lines else None)
    if use is None:
        raise LatticeFileError("no line to use")
    names = P.expand(use)
    if P.start_at is not None:
        # start_branch_at: the elements before the named one move to the end of the branch
        if P.start_at not in names:
            raise LatticeFileError(f"start_branch_at {P.start_at!r}: no such element in the line")
        i = names.index(P.start_at)
        names = names[i:] + names[:i]
    P.apply_overlays()
    eles = []
    # total length, for rf cavities given by harmonic number (f = harmon beta0 c / L_total)
    def arc_length(a, tname):
        L, ang = a.get("l", 0.0), a.get("angle")
        if tname == "rbend" and ang not in (None, 0.0):
            return L * ang / (2 * math.sin(ang / 2))
        return L
    l_total = sum(arc_length(P.eles[nm][1], P.eles[nm][0]) for nm in names)
    for nm in names:
        tname, attrs = P.eles[nm]
        a = dict(attrs)
        default_tilt = a.pop("__default_tilt__", False)
        L = a.get("l", 0.0)
        if tname in ("sbend", "rbend"):
            angle = a.pop("angle", None)
            rho = a.pop("rho", None)
            if tname == "rbend" and angle not in (None, 0.0):
                # rbend -> sbend: l is the chord; the edges turn by half the bend angle
                L = L * angle / (2 * math.sin(angle / 2))
                a["l"] = L
                a["e1"] = a.get("e1", 0.0) + angle / 2
                a["e2"] = a.get("e2", 0.0) + angle / 2
            if angle is not None and L != 0:
                a["g"] = angle / L
            elif rho is not None:
                a["g"] = 1.0 / rho
        else:
            for k in ("angle", "rho"):
                if k in a:
                    raise LatticeFileError(f"{k} on a {tname}")
        harmon = a.pop("harmon", None)
        if harmon is not None and "rf_frequency" not in a:
            if tname != "rfcavity" or l_total <= 0:
                raise LatticeFileError("harmon needs an rf cavity in a line of nonzero length")
            a["rf_frequency"] = harmon * (p0c / e_tot) * E.C_LIGHT / l_total
        num_steps = int(a.pop("num_steps", 0))
        ds_step = a.pop("ds_step", 0.0)
        poles = {"a_pole": {}, "b_pole": {}, "a_pole_elec": {}, "b_pole_elec": {}}
        for k in [k for k in a if _MULTIPOLE.match(k)]:
            ab, n, elec = _MULTIPOLE.match(k).groups()
            n = int(n)
            if n >= E.N_POLE:
                raise LatticeFileError(f"multipole order {n} above {E.N_POLE - 1}")
            poles[f"{ab}_pole{'_elec' if elec else ''}"][n] = a.pop(k)
        poles = {k: v for k, v in poles.items() if v}
        thin = {"knl": {}, "tn": {}, "ksl": {}}
        for k in [k for k in a if _THIN.match(k)]:
            kn, skew, tn = _THIN.match(k).groups()
            n = int(kn if kn is not None else tn)
            if tname != "multipole":
                raise LatticeFileError(f"{k} on a {tname} (only a multipole has it)")
            if n >= E.N_POLE:
                raise LatticeFileError(f"multipole order {n} above {E.N_POLE - 1}")
            thin["tn" if tn is not None else ("ksl" if skew else "knl")][n] = a.pop(k)
        if tname == "multipole":
            if poles:
                raise LatticeFileError("a multipole takes KnL, KnSL and Tn, not a/b values")
            poles = {k: v for k, v in thin.items() if v}
            if "knl" not in poles and poles:
                poles["knl"] = {}
        eles.append(E.make_element(_TYPES[tname], p0c, e_tot, num_steps=num_steps, ds_step=ds_step,
                                   default_tilt=default_tilt, **poles, **a))
    return eles, mc2, charge
