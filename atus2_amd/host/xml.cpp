// xml.cpp -- see xml.h.
#include "xml.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace atus_host {

const std::string* XmlNode::attr(const std::string& key) const {
  for (const auto& kv : attrs)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

std::string XmlNode::child_value() const {
  for (const auto& c : children)
    if (c->is_text()) return c->text;
  return std::string();
}

bool XmlNode::has_text_child() const {
  for (const auto& c : children)
    if (c->is_text()) return true;
  return false;
}

void XmlNode::descendants(std::vector<const XmlNode*>& out) const {
  for (const auto& c : children) {
    if (c->is_text()) continue;
    out.push_back(c.get());
    c->descendants(out);
  }
}

std::vector<const XmlNode*> XmlNode::elements() const {
  std::vector<const XmlNode*> out;
  for (const auto& c : children)
    if (!c->is_text()) out.push_back(c.get());
  return out;
}

namespace {

class XmlParser {
 public:
  explicit XmlParser(const std::string& s) : m_s(s) {}

  std::unique_ptr<XmlNode> parse_document() {
    auto doc = std::make_unique<XmlNode>();
    doc->name = "#document";
    parse_content(*doc, true);
    return doc;
  }

 private:
  const std::string& m_s;
  size_t m_p = 0;

  [[noreturn]] void error(const std::string& what) {
    size_t line = 1;
    for (size_t i = 0; i < m_p && i < m_s.size(); ++i)
      if (m_s[i] == '\n') ++line;
    throw std::string("XML parse error at line " + std::to_string(line) + ": " + what);
  }
  bool starts(const char* t) const { return m_s.compare(m_p, std::char_traits<char>::length(t), t) == 0; }
  void skip_to(const char* t) {
    size_t e = m_s.find(t, m_p);
    if (e == std::string::npos) error(std::string("unterminated construct, expected ") + t);
    m_p = e + std::char_traits<char>::length(t);
  }
  static bool name_char(char c) {
    return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':';
  }
  std::string parse_name() {
    size_t b = m_p;
    while (m_p < m_s.size() && name_char(m_s[m_p])) ++m_p;
    if (b == m_p) error("expected a name");
    return m_s.substr(b, m_p - b);
  }
  void skip_ws() {
    while (m_p < m_s.size() && std::isspace((unsigned char)m_s[m_p])) ++m_p;
  }
  std::string decode(const std::string& raw) {
    std::string out;
    out.reserve(raw.size());
    for (size_t i = 0; i < raw.size(); ++i) {
      if (raw[i] != '&') {
        out += raw[i];
        continue;
      }
      size_t e = raw.find(';', i);
      if (e == std::string::npos) {
        out += raw[i];
        continue;
      }
      std::string ent = raw.substr(i + 1, e - i - 1);
      if (ent == "lt") out += '<';
      else if (ent == "gt") out += '>';
      else if (ent == "amp") out += '&';
      else if (ent == "quot") out += '"';
      else if (ent == "apos") out += '\'';
      else if (!ent.empty() && ent[0] == '#') {
        long v = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                   : std::strtol(ent.c_str() + 1, nullptr, 10);
        if (v > 0 && v < 128) out += (char)v;
      } else {
        out += raw.substr(i, e - i + 1);
      }
      i = e;
    }
    return out;
  }
  void add_text(XmlNode& parent, const std::string& raw, bool keep_ws_only) {
    if (raw.empty()) return;
    bool all_ws = true;
    for (char c : raw)
      if (!std::isspace((unsigned char)c)) all_ws = false;
    // whitespace-only runs between elements are dropped, as pugixml does by default
    if (all_ws && !keep_ws_only) return;
    auto t = std::make_unique<XmlNode>();
    t->text = decode(raw);
    parent.children.push_back(std::move(t));
  }
  // parses children of `parent` until its end tag (or end of input at top level)
  void parse_content(XmlNode& parent, bool top) {
    while (m_p < m_s.size()) {
      if (m_s[m_p] != '<') {
        size_t b = m_p;
        while (m_p < m_s.size() && m_s[m_p] != '<') ++m_p;
        if (!top) add_text(parent, m_s.substr(b, m_p - b), false);
        continue;
      }
      if (starts("<!--")) {
        skip_to("-->");
      } else if (starts("<![CDATA[")) {
        m_p += 9;
        size_t e = m_s.find("]]>", m_p);
        if (e == std::string::npos) error("unterminated CDATA");
        if (!top) add_text(parent, m_s.substr(m_p, e - m_p), true);
        m_p = e + 3;
      } else if (starts("<?")) {
        skip_to("?>");
      } else if (starts("<!")) {
        skip_to(">");
      } else if (starts("</")) {
        if (top) error("unexpected end tag");
        m_p += 2;
        std::string n = parse_name();
        if (n != parent.name) error("end tag </" + n + "> does not match <" + parent.name + ">");
        skip_ws();
        if (m_p >= m_s.size() || m_s[m_p] != '>') error("expected '>'");
        ++m_p;
        return;
      } else {
        ++m_p;
        auto el = std::make_unique<XmlNode>();
        el->name = parse_name();
        for (;;) {
          skip_ws();
          if (m_p >= m_s.size()) error("unterminated start tag");
          if (m_s[m_p] == '/' || m_s[m_p] == '>') break;
          std::string an = parse_name();
          skip_ws();
          if (m_p >= m_s.size() || m_s[m_p] != '=') error("expected '=' after attribute " + an);
          ++m_p;
          skip_ws();
          if (m_p >= m_s.size() || (m_s[m_p] != '"' && m_s[m_p] != '\'')) error("expected quoted value");
          char q = m_s[m_p++];
          size_t e = m_s.find(q, m_p);
          if (e == std::string::npos) error("unterminated attribute value");
          el->attrs.emplace_back(an, decode(m_s.substr(m_p, e - m_p)));
          m_p = e + 1;
        }
        bool self_close = m_s[m_p] == '/';
        if (self_close) {
          ++m_p;
          if (m_p >= m_s.size() || m_s[m_p] != '>') error("expected '>'");
        }
        ++m_p;
        XmlNode& ref = *el;
        parent.children.push_back(std::move(el));
        if (!self_close) parse_content(ref, false);
      }
    }
    if (!top) error("missing end tag </" + parent.name + ">");
  }
};

}  // namespace

std::unique_ptr<XmlNode> xml_parse_string(const std::string& text) {
  XmlParser p(text);
  return p.parse_document();
}

std::unique_ptr<XmlNode> xml_parse_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) throw std::string("Could not open file " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  std::string s = ss.str();
  return xml_parse_string(s);
}

}  // namespace atus_host
