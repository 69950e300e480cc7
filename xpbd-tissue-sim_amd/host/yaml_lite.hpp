// A small reader for the YAML subset the simulation configs use: block
// mappings and sequences by indentation, inline lists [a, b, c], plain and
// quoted scalars, '#' comments.  Anchors, multi-line scalars and flow
// mappings are not supported and are reported as errors.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace yaml_lite {

struct Node {
    enum Kind { Null, Scalar, Map, Seq } kind = Null;
    std::string scalar;
    std::vector<std::pair<std::string, Node>> map;  // insertion order kept
    std::vector<Node> seq;

    bool has(const std::string& key) const;
    const Node& operator[](const std::string& key) const;  // throws if missing
    const Node* find(const std::string& key) const;
    double as_double() const;
    long as_int() const;
    bool as_bool() const;
    const std::string& as_string() const;
    std::vector<double> as_doubles() const;
};

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

Node parse(const std::string& text);
Node parse_file(const std::string& path);

}  // namespace yaml_lite
