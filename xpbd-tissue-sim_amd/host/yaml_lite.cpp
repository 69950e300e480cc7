#include "yaml_lite.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace yaml_lite {

namespace {

struct Line {
    int indent;
    std::string text;  // without indentation and comment
    int number;
};

std::string strip_comment(const std::string& s)
{
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); i++) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
    }
    return s;
}

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string unquote(const std::string& s)
{
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

// position of the "key:" separator (a colon followed by space or end, outside quotes)
size_t key_colon(const std::string& s)
{
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); i++) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == ':' && !sq && !dq && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
    }
    return std::string::npos;
}

Node scalar_or_inline(const std::string& raw, int line)
{
    const std::string s = trim(raw);
    Node n;
    if (s.empty()) return n;
    if (s.front() == '[') {
        if (s.back() != ']') throw ParseError("line " + std::to_string(line) + ": unterminated inline list");
        n.kind = Node::Seq;
        const std::string body = s.substr(1, s.size() - 2);
        std::string cur;
        bool sq = false, dq = false;
        for (char c : body) {
            if (c == '\'' && !dq) sq = !sq;
            if (c == '"' && !sq) dq = !dq;
            if (c == ',' && !sq && !dq) {
                n.seq.push_back(scalar_or_inline(cur, line));
                cur.clear();
            } else {
                cur += c;
            }
        }
        if (!trim(cur).empty()) n.seq.push_back(scalar_or_inline(cur, line));
        return n;
    }
    if (s.front() == '{') {
        // flat flow mapping {k: v, k2: [a, b]}
        if (s.back() != '}') throw ParseError("line " + std::to_string(line) + ": unterminated inline mapping");
        n.kind = Node::Map;
        const std::string body = s.substr(1, s.size() - 2);
        std::vector<std::string> items;
        std::string cur;
        bool sq = false, dq = false;
        int depth = 0;
        for (char c : body) {
            if (c == '\'' && !dq) sq = !sq;
            if (c == '"' && !sq) dq = !dq;
            if (!sq && !dq && (c == '[' || c == '{')) depth++;
            if (!sq && !dq && (c == ']' || c == '}')) depth--;
            if (c == ',' && !sq && !dq && depth == 0) {
                items.push_back(cur);
                cur.clear();
            } else {
                cur += c;
            }
        }
        if (!trim(cur).empty()) items.push_back(cur);
        for (const std::string& it : items) {
            const std::string t = trim(it);
            const size_t c = key_colon(t);
            if (c == std::string::npos)
                throw ParseError("line " + std::to_string(line) + ": expected 'key: value' in '" + t + "'");
            n.map.emplace_back(unquote(trim(t.substr(0, c))), scalar_or_inline(t.substr(c + 1), line));
        }
        return n;
    }
    if (s.front() == '&' || s.front() == '*' || s.front() == '|' || s.front() == '>')
        throw ParseError("line " + std::to_string(line) + ": unsupported YAML construct '" + s + "'");
    n.kind = Node::Scalar;
    n.scalar = unquote(s);
    return n;
}

class Parser {
public:
    explicit Parser(std::vector<Line> lines) : L(std::move(lines)) {}
    Node parse_block(size_t& i, int indent);

private:
    std::vector<Line> L;
};

Node Parser::parse_block(size_t& i, int indent)
{
    Node n;
    if (i >= L.size()) return n;
    const bool seq = L[i].text.rfind("- ", 0) == 0 || L[i].text == "-";
    n.kind = seq ? Node::Seq : Node::Map;
    while (i < L.size() && L[i].indent == indent) {
        const Line& ln = L[i];
        if (seq) {
            if (!(ln.text.rfind("- ", 0) == 0 || ln.text == "-"))
                throw ParseError("line " + std::to_string(ln.number) + ": expected a list item");
            const std::string rest = ln.text == "-" ? "" : trim(ln.text.substr(2));
            i++;
            if (rest.empty()) {
                if (i < L.size() && L[i].indent > indent) n.seq.push_back(parse_block(i, L[i].indent));
                else n.seq.push_back(Node());
            } else if (key_colon(rest) != std::string::npos && rest.front() != '[' && rest.front() != '{') {
                // "- key: value" starts a mapping whose further keys are indented past the dash
                const int item_indent = indent + 2;
                std::vector<Line> sub{ { item_indent, rest, ln.number } };
                Node item;
                size_t j = 0;
                // collect the continuation lines of this item
                while (i < L.size() && L[i].indent > indent) sub.push_back(L[i++]);
                // re-base indentation so the first key and the continuation align
                int cont = item_indent;
                for (size_t k = 1; k < sub.size(); k++)
                    if (sub[k].indent < cont) cont = sub[k].indent;
                sub[0].indent = cont;
                Parser p(sub);
                item = p.parse_block(j, cont);
                n.seq.push_back(item);
            } else {
                n.seq.push_back(scalar_or_inline(rest, ln.number));
            }
        } else {
            const size_t c = key_colon(ln.text);
            if (c == std::string::npos)
                throw ParseError("line " + std::to_string(ln.number) + ": expected 'key: value'");
            const std::string key = unquote(trim(ln.text.substr(0, c)));
            const std::string val = trim(ln.text.substr(c + 1));
            i++;
            Node child;
            if (val.empty()) {
                if (i < L.size() && L[i].indent > indent) child = parse_block(i, L[i].indent);
                else if (i < L.size() && L[i].indent == indent && L[i].text.rfind("- ", 0) == 0)
                    child = parse_block(i, indent);  // a sequence at the key's own indentation
            } else {
                child = scalar_or_inline(val, ln.number);
            }
            for (auto& kv : n.map)
                if (kv.first == key) throw ParseError("line " + std::to_string(ln.number) + ": duplicate key " + key);
            n.map.emplace_back(key, std::move(child));
        }
    }
    if (i < L.size() && L[i].indent > indent)
        throw ParseError("line " + std::to_string(L[i].number) + ": unexpected indentation");
    return n;
}

}  // namespace

bool Node::has(const std::string& key) const { return find(key) != nullptr; }

const Node* Node::find(const std::string& key) const
{
    if (kind != Map) return nullptr;
    for (const auto& kv : map)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

const Node& Node::operator[](const std::string& key) const
{
    const Node* n = find(key);
    if (!n) throw ParseError("missing key '" + key + "'");
    return *n;
}

const std::string& Node::as_string() const
{
    if (kind != Scalar) throw ParseError("expected a scalar");
    return scalar;
}

double Node::as_double() const
{
    const std::string& s = as_string();
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0') throw ParseError("not a number: '" + s + "'");
    return v;
}

long Node::as_int() const
{
    const double v = as_double();
    const long i = (long)v;
    if ((double)i != v) throw ParseError("not an integer: '" + scalar + "'");
    return i;
}

bool Node::as_bool() const
{
    const std::string& s = as_string();
    if (s == "true" || s == "True" || s == "yes") return true;
    if (s == "false" || s == "False" || s == "no") return false;
    throw ParseError("not a boolean: '" + s + "'");
}

std::vector<double> Node::as_doubles() const
{
    if (kind != Seq) throw ParseError("expected a list");
    std::vector<double> out;
    for (const Node& n : seq) out.push_back(n.as_double());
    return out;
}

Node parse(const std::string& text)
{
    std::vector<Line> lines;
    std::istringstream is(text);
    std::string raw;
    int no = 0;
    while (std::getline(is, raw)) {
        no++;
        if (raw.find('\t') != std::string::npos) {
            // tabs are not YAML indentation
            const size_t first = raw.find_first_not_of(" \t");
            if (first != std::string::npos && raw.substr(0, first).find('\t') != std::string::npos)
                throw ParseError("line " + std::to_string(no) + ": tab in indentation");
        }
        const std::string s = strip_comment(raw);
        const std::string t = trim(s);
        if (t.empty() || t == "---") continue;
        int ind = 0;
        while (ind < (int)s.size() && s[ind] == ' ') ind++;
        lines.push_back({ ind, t, no });
    }
    if (lines.empty()) return Node();
    size_t i = 0;
    Parser p(lines);
    Node root = p.parse_block(i, lines[0].indent);
    if (i != lines.size()) throw ParseError("line " + std::to_string(lines[i].number) + ": bad indentation");
    return root;
}

Node parse_file(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw ParseError("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace yaml_lite
