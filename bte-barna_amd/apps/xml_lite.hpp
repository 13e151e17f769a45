// Minimal XML reader for the solver inputs: elements, attributes and text.
// Handles the prolog, comments, self-closing tags and the five predefined
// entities; no DTDs or namespaces. Lookup follows parseXMLfield
// (include/io_utils.hpp:184-208): a field is an attribute of the element or
// the text of a child element of that name.
#pragma once
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace xml {

struct Node {
    std::string name, text;
    std::map<std::string, std::string> attr;
    std::vector<std::unique_ptr<Node>> kids;

    const Node* child(const std::string& n) const {
        for (auto& k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
    std::vector<const Node*> children(const std::string& n) const {
        std::vector<const Node*> out;
        for (auto& k : kids)
            if (k->name == n) out.push_back(k.get());
        return out;
    }
    bool has(const std::string& f) const { return attr.count(f) || child(f); }
    std::string str(const std::string& f) const {
        auto it = attr.find(f);
        if (it != attr.end()) return it->second;
        if (auto* c = child(f)) return trim(c->text);
        throw std::runtime_error("<" + name + ">: missing field '" + f + "'");
    }
    std::string str(const std::string& f, const std::string& dflt) const { return has(f) ? str(f) : dflt; }
    double num(const std::string& f) const {
        std::string s = str(f);
        char* end = nullptr;
        double v = std::strtod(s.c_str(), &end);
        if (end == s.c_str()) throw std::runtime_error("<" + name + ">: field '" + f + "' is not a number: " + s);
        return v;
    }
    double num(const std::string& f, double dflt) const { return has(f) ? num(f) : dflt; }
    static std::string trim(const std::string& s) {
        size_t a = 0, b = s.size();
        while (a < b && std::isspace((unsigned char)s[a])) a++;
        while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
        return s.substr(a, b - a);
    }
};

class Parser {
public:
    explicit Parser(std::string src) : s_(std::move(src)) {}
    std::unique_ptr<Node> parse() {
        skip_misc();
        auto root = element();
        skip_misc();
        if (p_ != s_.size()) fail("content after the root element");
        return root;
    }

private:
    std::string s_;
    size_t p_ = 0;
    [[noreturn]] void fail(const std::string& what) {
        size_t line = 1;
        for (size_t i = 0; i < p_ && i < s_.size(); i++) line += s_[i] == '\n';
        throw std::runtime_error("XML parse error at line " + std::to_string(line) + ": " + what);
    }
    bool starts(const char* t) const { return s_.compare(p_, std::char_traits<char>::length(t), t) == 0; }
    void ws() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) p_++;
    }
    void skip_misc() {
        for (;;) {
            ws();
            if (starts("<?")) {
                size_t e = s_.find("?>", p_);
                if (e == std::string::npos) fail("unterminated prolog");
                p_ = e + 2;
            } else if (starts("<!--")) {
                size_t e = s_.find("-->", p_);
                if (e == std::string::npos) fail("unterminated comment");
                p_ = e + 3;
            } else if (starts("<!")) {
                size_t e = s_.find('>', p_);
                if (e == std::string::npos) fail("unterminated declaration");
                p_ = e + 1;
            } else {
                return;
            }
        }
    }
    std::string name() {
        size_t a = p_;
        while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_' || s_[p_] == '-' ||
                                  s_[p_] == '.' || s_[p_] == ':'))
            p_++;
        if (a == p_) fail("expected a name");
        return s_.substr(a, p_ - a);
    }
    static std::string unescape(const std::string& v) {
        std::string o;
        for (size_t i = 0; i < v.size(); i++) {
            if (v[i] != '&') { o += v[i]; continue; }
            size_t e = v.find(';', i);
            std::string ent = e == std::string::npos ? "" : v.substr(i + 1, e - i - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else { o += '&'; continue; }
            i = e;
        }
        return o;
    }
    std::unique_ptr<Node> element() {
        if (!starts("<")) fail("expected '<'");
        p_++;
        auto n = std::make_unique<Node>();
        n->name = name();
        for (;;) {
            ws();
            if (starts("/>")) { p_ += 2; return n; }
            if (starts(">")) { p_++; break; }
            std::string k = name();
            ws();
            if (!starts("=")) fail("expected '=' after attribute " + k);
            p_++;
            ws();
            if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) fail("expected a quoted value");
            char q = s_[p_++];
            size_t e = s_.find(q, p_);
            if (e == std::string::npos) fail("unterminated attribute value");
            n->attr[k] = unescape(s_.substr(p_, e - p_));
            p_ = e + 1;
        }
        for (;;) {
            if (p_ >= s_.size()) fail("unterminated element <" + n->name + ">");
            if (starts("<!--")) {
                size_t e = s_.find("-->", p_);
                if (e == std::string::npos) fail("unterminated comment");
                p_ = e + 3;
            } else if (starts("</")) {
                p_ += 2;
                std::string c = name();
                if (c != n->name) fail("mismatched </" + c + "> for <" + n->name + ">");
                ws();
                if (!starts(">")) fail("expected '>'");
                p_++;
                return n;
            } else if (starts("<")) {
                n->kids.push_back(element());
            } else {
                size_t e = s_.find('<', p_);
                if (e == std::string::npos) fail("unterminated element <" + n->name + ">");
                n->text += unescape(s_.substr(p_, e - p_));
                p_ = e;
            }
        }
    }
};

inline std::unique_ptr<Node> parse_file(const std::string& fn) {
    std::ifstream f(fn);
    if (!f) throw std::runtime_error("cannot open " + fn);
    std::stringstream ss;
    ss << f.rdbuf();
    return Parser(ss.str()).parse();
}

}  // namespace xml
