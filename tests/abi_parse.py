"""Parse C struct/union declarations into ordered (type, name, dims) member lists.

Used by the ABI layout test to compare include/shape_abi.h with the member lists
of the reference's src/shape/shape2.h (read as text, never compiled).  Handles
what those two headers contain: nested unions, comma-separated declarators,
pointer stars, array dimensions and // or /* */ comments.
"""
import re


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return text


def _expand_macros(text):
    """Substitute object-like integer #defines used as array sizes."""
    defs = dict(re.findall(r"#define\s+(\w+)\s+(\d+)\s*$", text, flags=re.M))
    text = re.sub(r"^\s*#.*$", " ", text, flags=re.M)

    def sub_dim(m):
        expr = m.group(1)
        for k, v in defs.items():
            expr = re.sub(r"\b%s\b" % k, v, expr)
        try:
            return "[%d]" % eval(expr, {"__builtins__": {}})
        except Exception:
            return "[%s]" % expr.strip()

    return re.sub(r"\[([^\]]*)\]", sub_dim, text)


def _split_members(body):
    """Split a struct body into top-level declarations (nested braces kept)."""
    out, depth, cur = [], 0, []
    for ch in body:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == ";" and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    return [d for d in out if d]


def _parse_decl(decl):
    """Return a list of (type, name, dims) for one declaration statement."""
    decl = " ".join(decl.split())
    m = re.match(r"^(union|struct)\s+(\w+)\s*\{(.*)\}\s*(\w+)$", decl, flags=re.S)
    if m:  # nested aggregate with a member name
        kind, tag, body, name = m.groups()
        inner = [x for d in _split_members(body) for x in _parse_decl(d)]
        return [("%s %s{%s}" % (kind, tag, ";".join("%s %s%s" % t for t in inner)), name, "")]
    m = re.match(r"^((?:unsigned\s+|signed\s+|const\s+)*(?:struct\s+\w+|union\s+\w+|\w+))\s+(.*)$", decl)
    if not m:
        raise ValueError("cannot parse declaration: %r" % decl)
    base, rest = m.groups()
    members = []
    for d in rest.split(","):
        d = d.strip()
        stars = len(d) - len(d.lstrip("*"))
        d = d.lstrip("*").strip()
        nm = re.match(r"^(\w+)\s*((?:\[[^\]]*\]\s*)*)$", d)
        if not nm:
            raise ValueError("cannot parse declarator: %r in %r" % (d, decl))
        name, dims = nm.groups()
        members.append((" ".join(base.split()) + " " + "*" * stars, name, dims.replace(" ", "")))
    return members


def parse_structs(text):
    """Map struct/union tag -> ordered member list, for every top-level definition."""
    text = _expand_macros(_strip_comments(text))
    result = {}
    for m in re.finditer(r"\b(struct|union)\s+(\w+)\s*\{", text):
        start = m.end()
        depth, i = 1, start
        while depth:
            if text[i] == "{":
                depth += 1
            elif text[i] == "}":
                depth -= 1
            i += 1
        body = text[start:i - 1]
        tag = m.group(2)
        if tag in result:
            continue
        result[tag] = [x for d in _split_members(body) for x in _parse_decl(d)]
    return result
