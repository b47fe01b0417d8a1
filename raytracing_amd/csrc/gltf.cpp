// glTF 2.0 loading (host code, gltf.hpp): a small JSON parser, the GLB container, buffers, accessors, triangle
// primitives (lists, strips and fans), the node hierarchy and metallic-roughness factors, after the glTF 2.0
// specification (Khronos, sections 3.6-3.9 and 4.4).
#include "gltf.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <variant>

namespace rt
{
namespace
{
// --- JSON -----------------------------------------------------------------------------------------------------------
struct json;
using json_object = std::map<std::string, json>;
using json_array = std::vector<json>;
struct json
{
    std::variant<std::nullptr_t, bool, double, std::string, std::shared_ptr<json_array>, std::shared_ptr<json_object>>
        v = nullptr;

    bool is_object() const { return std::holds_alternative<std::shared_ptr<json_object>>(v); }
    bool is_array() const { return std::holds_alternative<std::shared_ptr<json_array>>(v); }
    bool is_number() const { return std::holds_alternative<double>(v); }
    bool is_string() const { return std::holds_alternative<std::string>(v); }
    const json_object& object() const
    {
        if (!is_object())
        {
            throw std::runtime_error("the glTF JSON has a value where an object belongs");
        }
        return *std::get<std::shared_ptr<json_object>>(v);
    }
    const json_array& array() const
    {
        if (!is_array())
        {
            throw std::runtime_error("the glTF JSON has a value where an array belongs");
        }
        return *std::get<std::shared_ptr<json_array>>(v);
    }
    double number() const
    {
        if (!is_number())
        {
            throw std::runtime_error("the glTF JSON has a value where a number belongs");
        }
        return std::get<double>(v);
    }
    const std::string& string() const
    {
        if (!is_string())
        {
            throw std::runtime_error("the glTF JSON has a value where a string belongs");
        }
        return std::get<std::string>(v);
    }
    // A member of an object, or null when it is absent
    const json* find(const std::string& key) const
    {
        const json_object& o = object();
        const auto it = o.find(key);
        return it == o.end() ? nullptr : &it->second;
    }
};

class json_parser
{
public:
    explicit json_parser(const std::string& text) : s_(text) {}
    json parse()
    {
        json j = value(0);
        skip();
        if (p_ != s_.size())
        {
            fail("text after the JSON");
        }
        return j;
    }

private:
    const std::string& s_;
    std::size_t p_ = 0;

    [[noreturn]] void fail(const std::string& what) const
    {
        throw std::runtime_error("the glTF JSON doesn't parse: " + what + " at byte " + std::to_string(p_));
    }
    void skip()
    {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r'))
        {
            ++p_;
        }
    }
    bool take(char c)
    {
        skip();
        if (p_ < s_.size() && s_[p_] == c)
        {
            ++p_;
            return true;
        }
        return false;
    }
    json value(int depth)
    {
        if (depth > 256)
        {
            fail("nesting too deep");
        }
        skip();
        if (p_ >= s_.size())
        {
            fail("the end of the text");
        }
        const char c = s_[p_];
        json j;
        if (c == '{')
        {
            ++p_;
            auto o = std::make_shared<json_object>();
            if (!take('}'))
            {
                do
                {
                    skip();
                    if (p_ >= s_.size() || s_[p_] != '"')
                    {
                        fail("a key that isn't a string");
                    }
                    std::string key = string_value();
                    if (!take(':'))
                    {
                        fail("a missing ':'");
                    }
                    (*o)[key] = value(depth + 1);
                } while (take(','));
                if (!take('}'))
                {
                    fail("a missing '}'");
                }
            }
            j.v = o;
        }
        else if (c == '[')
        {
            ++p_;
            auto a = std::make_shared<json_array>();
            if (!take(']'))
            {
                do
                {
                    a->push_back(value(depth + 1));
                } while (take(','));
                if (!take(']'))
                {
                    fail("a missing ']'");
                }
            }
            j.v = a;
        }
        else if (c == '"')
        {
            j.v = string_value();
        }
        else if (s_.compare(p_, 4, "true") == 0)
        {
            p_ += 4;
            j.v = true;
        }
        else if (s_.compare(p_, 5, "false") == 0)
        {
            p_ += 5;
            j.v = false;
        }
        else if (s_.compare(p_, 4, "null") == 0)
        {
            p_ += 4;
        }
        else
        {
            const char* begin = s_.c_str() + p_;
            char* end = nullptr;
            const double x = std::strtod(begin, &end);
            if (end == begin || !(c == '-' || (c >= '0' && c <= '9')))
            {
                fail("an unexpected character");
            }
            p_ += static_cast<std::size_t>(end - begin);
            j.v = x;
        }
        return j;
    }
    static void append_utf8(std::string& out, uint32_t cp)
    {
        if (cp < 0x80)
        {
            out += static_cast<char>(cp);
        }
        else if (cp < 0x800)
        {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
        else if (cp < 0x10000)
        {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
        else
        {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4()
    {
        if (p_ + 4 > s_.size())
        {
            fail("a short \\u escape");
        }
        uint32_t cp = 0;
        for (int k = 0; k < 4; ++k)
        {
            const char h = s_[p_++];
            cp <<= 4;
            if (h >= '0' && h <= '9')
            {
                cp |= static_cast<uint32_t>(h - '0');
            }
            else if (h >= 'a' && h <= 'f')
            {
                cp |= static_cast<uint32_t>(h - 'a' + 10);
            }
            else if (h >= 'A' && h <= 'F')
            {
                cp |= static_cast<uint32_t>(h - 'A' + 10);
            }
            else
            {
                fail("a bad \\u escape");
            }
        }
        return cp;
    }
    std::string string_value()
    {
        ++p_; // the opening quote
        std::string out;
        while (true)
        {
            if (p_ >= s_.size())
            {
                fail("a string without its end");
            }
            const char c = s_[p_++];
            if (c == '"')
            {
                return out;
            }
            if (c != '\\')
            {
                out += c;
                continue;
            }
            if (p_ >= s_.size())
            {
                fail("a string without its end");
            }
            const char e = s_[p_++];
            switch (e)
            {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u':
            {
                uint32_t cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && p_ + 1 < s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u')
                {
                    p_ += 2;
                    const uint32_t low = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (low - 0xDC00);
                }
                append_utf8(out, cp);
                break;
            }
            default: fail("a bad escape");
            }
        }
    }
};

// --- files and buffers ------------------------------------------------------------------------------------------------
std::vector<uint8_t> read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
    {
        throw std::runtime_error("can't open " + path);
    }
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

uint32_t le32(const std::vector<uint8_t>& b, std::size_t at)
{
    return static_cast<uint32_t>(b[at]) | static_cast<uint32_t>(b[at + 1]) << 8 | static_cast<uint32_t>(b[at + 2]) << 16 |
           static_cast<uint32_t>(b[at + 3]) << 24;
}

std::vector<uint8_t> base64(const std::string& text, std::size_t from)
{
    auto value = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    std::vector<uint8_t> out;
    uint32_t bits = 0;
    int count = 0;
    for (std::size_t k = from; k < text.size(); ++k)
    {
        if (text[k] == '=')
        {
            break;
        }
        const int v = value(text[k]);
        if (v < 0)
        {
            throw std::runtime_error("a data: URI that isn't base64");
        }
        bits = bits << 6 | static_cast<uint32_t>(v);
        count += 6;
        if (count >= 8)
        {
            count -= 8;
            out.push_back(static_cast<uint8_t>(bits >> count));
        }
    }
    return out;
}

std::string percent_decoded(const std::string& uri)
{
    std::string out;
    for (std::size_t k = 0; k < uri.size(); ++k)
    {
        if (uri[k] == '%' && k + 2 < uri.size())
        {
            out += static_cast<char>(std::stoi(uri.substr(k + 1, 2), nullptr, 16));
            k += 2;
        }
        else
        {
            out += uri[k];
        }
    }
    return out;
}

// The most elements one accessor may hold: 2^28 (a mesh of 268 M vertices or 89 M triangles), far beyond any model
// this renderer's memory holds, and small enough that element counts times components or strides never wrap
constexpr std::size_t max_accessor_elements = std::size_t{1} << 28;

std::size_t index_of(const json* j, const char* what)
{
    if (!j)
    {
        throw std::runtime_error(std::string("a missing ") + what);
    }
    const double x = j->number();
    if (!(x >= 0) || x != std::floor(x) || x > 1e15)
    {
        throw std::runtime_error(std::string("a bad ") + what);
    }
    return static_cast<std::size_t>(x);
}

double number_or(const json& o, const char* key, double fallback)
{
    const json* j = o.find(key);
    return j ? j->number() : fallback;
}

const json_array& array_or_empty(const json& root, const char* key)
{
    static const json_array empty;
    const json* j = root.find(key);
    return j ? j->array() : empty;
}

// 4x4 matrices, column-major as glTF stores them
using mat4 = std::array<double, 16>;
mat4 identity()
{
    mat4 m{};
    m[0] = m[5] = m[10] = m[15] = 1;
    return m;
}
mat4 multiply(const mat4& a, const mat4& b)
{
    mat4 c{};
    for (int col = 0; col < 4; ++col)
    {
        for (int row = 0; row < 4; ++row)
        {
            double s = 0;
            for (int k = 0; k < 4; ++k)
            {
                s += a[k * 4 + row] * b[col * 4 + k];
            }
            c[col * 4 + row] = s;
        }
    }
    return c;
}
// A node's local transform: its matrix, or translation * rotation * scale
mat4 local_transform(const json& node)
{
    if (const json* m = node.find("matrix"))
    {
        const json_array& a = m->array();
        if (a.size() != 16)
        {
            throw std::runtime_error("a node's matrix without 16 numbers");
        }
        mat4 out;
        for (int k = 0; k < 16; ++k)
        {
            out[k] = a[k].number();
        }
        return out;
    }
    double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
    if (const json* j = node.find("translation"))
    {
        for (int k = 0; k < 3; ++k)
        {
            t[k] = j->array().at(k).number();
        }
    }
    if (const json* j = node.find("rotation"))
    {
        for (int k = 0; k < 4; ++k)
        {
            q[k] = j->array().at(k).number();
        }
    }
    if (const json* j = node.find("scale"))
    {
        for (int k = 0; k < 3; ++k)
        {
            s[k] = j->array().at(k).number();
        }
    }
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    // the rotation matrix of the unit quaternion (x, y, z, w), columns scaled
    const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w),     2 * (x * z - y * w),
                         2 * (x * y - z * w),     1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                         2 * (x * z + y * w),     2 * (y * z - x * w),     1 - 2 * (x * x + y * y)};
    mat4 m = identity();
    for (int col = 0; col < 3; ++col)
    {
        for (int row = 0; row < 3; ++row)
        {
            m[col * 4 + row] = r[col * 3 + row] * s[col];
        }
    }
    m[12] = t[0];
    m[13] = t[1];
    m[14] = t[2];
    return m;
}

class loader
{
public:
    explicit loader(const std::string& path) : path_(path)
    {
        const std::size_t slash = path.find_last_of("/\\");
        folder_ = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    }

    gltf_model load()
    {
        const std::vector<uint8_t> bytes = read_file(path_);
        std::string text;
        if (bytes.size() >= 4 && std::memcmp(bytes.data(), "glTF", 4) == 0)
        {
            read_glb(bytes, text);
        }
        else
        {
            text.assign(bytes.begin(), bytes.end());
        }
        root_ = json_parser(text).parse();
        if (!root_.is_object())
        {
            throw std::runtime_error("the glTF JSON isn't an object");
        }
        const json* asset = root_.find("asset");
        const json* version = asset ? asset->find("version") : nullptr;
        if (!version || version->string().rfind("2.", 0) != 0)
        {
            throw std::runtime_error("not a glTF 2.0 file (asset.version must be 2.x)");
        }
        for (const json& e : array_or_empty(root_, "extensionsRequired"))
        {
            throw std::runtime_error("it requires " + e.string() +
                                     ", which this renderer doesn't read (compressed meshes among such extensions): "
                                     "export the model without it");
        }
        load_buffers();
        load_materials();
        place_scene();
        return std::move(model_);
    }

private:
    std::string path_, folder_;
    json root_;
    std::vector<uint8_t> bin_; // a .glb's BIN chunk
    bool has_bin_ = false;
    std::vector<std::vector<uint8_t>> buffers_;
    gltf_model model_;
    std::map<std::pair<std::size_t, std::size_t>, int> primitive_mesh_; // (mesh, primitive) -> gltf_model mesh, -1 none
    bool warned_textures_ = false;

    // the binary container (glTF 2.0 section 4.4): a 12-byte header, a JSON chunk, then an optional BIN chunk
    void read_glb(const std::vector<uint8_t>& b, std::string& text)
    {
        if (b.size() < 12)
        {
            throw std::runtime_error("the file is shorter than it says (a .glb's header is 12 bytes)");
        }
        if (le32(b, 4) != 2)
        {
            throw std::runtime_error("a .glb of version " + std::to_string(le32(b, 4)) + ", not 2");
        }
        const std::size_t length = le32(b, 8);
        if (length > b.size())
        {
            throw std::runtime_error("the file is shorter than it says (" + std::to_string(b.size()) + " bytes of " +
                                     std::to_string(length) + ")");
        }
        std::size_t at = 12;
        bool json_seen = false;
        while (at + 8 <= length)
        {
            const std::size_t size = le32(b, at);
            const uint32_t type = le32(b, at + 4);
            if (at + 8 + size > length)
            {
                throw std::runtime_error("the file is shorter than it says (a chunk runs past its end)");
            }
            if (type == 0x4E4F534Au && !json_seen) // "JSON"
            {
                text.assign(b.begin() + static_cast<std::ptrdiff_t>(at + 8),
                            b.begin() + static_cast<std::ptrdiff_t>(at + 8 + size));
                json_seen = true;
            }
            else if (type == 0x004E4942u && json_seen && !has_bin_) // "BIN\0"
            {
                bin_.assign(b.begin() + static_cast<std::ptrdiff_t>(at + 8),
                            b.begin() + static_cast<std::ptrdiff_t>(at + 8 + size));
                has_bin_ = true;
            }
            at += 8 + ((size + 3) & ~std::size_t(3));
        }
        if (!json_seen)
        {
            throw std::runtime_error("a .glb without its JSON chunk");
        }
    }

    void load_buffers()
    {
        const json_array& buffers = array_or_empty(root_, "buffers");
        for (std::size_t k = 0; k < buffers.size(); ++k)
        {
            const json& b = buffers[k];
            const std::size_t length = index_of(b.find("byteLength"), "buffer byteLength");
            std::vector<uint8_t> data;
            if (const json* uri = b.find("uri"))
            {
                const std::string& u = uri->string();
                if (u.rfind("data:", 0) == 0)
                {
                    const std::size_t comma = u.find(',');
                    if (comma == std::string::npos || u.substr(0, comma).find(";base64") == std::string::npos)
                    {
                        throw std::runtime_error("a data: URI that isn't base64");
                    }
                    data = base64(u, comma + 1);
                }
                else
                {
                    data = read_file(folder_ + percent_decoded(u));
                }
            }
            else if (k == 0 && has_bin_)
            {
                data = bin_;
            }
            else
            {
                throw std::runtime_error("buffer " + std::to_string(k) + " has no data");
            }
            if (data.size() < length)
            {
                throw std::runtime_error("buffer " + std::to_string(k) + " is shorter than its byteLength");
            }
            buffers_.push_back(std::move(data));
        }
    }

    void load_materials()
    {
        for (const json& m : array_or_empty(root_, "materials"))
        {
            gltf_material out;
            if (const json* pbr = m.find("pbrMetallicRoughness"))
            {
                if (const json* c = pbr->find("baseColorFactor"))
                {
                    for (int k = 0; k < 3; ++k)
                    {
                        out.base_color[static_cast<std::size_t>(k)] = c->array().at(static_cast<std::size_t>(k)).number();
                    }
                }
                out.metallic = std::clamp(number_or(*pbr, "metallicFactor", 1.0), 0.0, 1.0);
                out.roughness = std::clamp(number_or(*pbr, "roughnessFactor", 1.0), 0.0, 1.0);
                if ((pbr->find("baseColorTexture") || pbr->find("metallicRoughnessTexture")) && !warned_textures_)
                {
                    model_.warnings.push_back("textures are not rendered: the materials' factors are used");
                    warned_textures_ = true;
                }
            }
            if (const json* e = m.find("emissiveFactor"))
            {
                double sum = 0;
                for (const json& x : e->array())
                {
                    sum += x.number();
                }
                if (sum > 0)
                {
                    model_.warnings.push_back("emissive materials are not rendered as lamps");
                }
            }
            model_.materials.push_back(out);
        }
    }

    // An accessor's elements as doubles (components per element), checked against its buffer view and buffer
    std::vector<double> read_accessor(std::size_t index, int& components, int& component_type)
    {
        const json_array& accessors = array_or_empty(root_, "accessors");
        if (index >= accessors.size())
        {
            throw std::runtime_error("an accessor index past the accessors");
        }
        const json& a = accessors[index];
        if (a.find("sparse"))
        {
            throw std::runtime_error("a sparse accessor, which this renderer doesn't read");
        }
        const std::string& type = a.find("type") ? a.find("type")->string() : std::string();
        components = type == "SCALAR" ? 1 : type == "VEC2" ? 2 : type == "VEC3" ? 3 : type == "VEC4" ? 4 : 0;
        if (components == 0)
        {
            throw std::runtime_error("an accessor of type '" + type + "'");
        }
        component_type = static_cast<int>(index_of(a.find("componentType"), "accessor componentType"));
        const std::size_t size = component_type == 5120 || component_type == 5121   ? 1
                                 : component_type == 5122 || component_type == 5123 ? 2
                                 : component_type == 5125 || component_type == 5126 ? 4
                                                                                    : 0;
        if (size == 0)
        {
            throw std::runtime_error("an accessor of component type " + std::to_string(component_type));
        }
        const std::size_t count = index_of(a.find("count"), "accessor count");
        // Every bound below is checked by subtraction or division, never by a sum or product that could wrap modulo
        // 2^64 (a crafted count or stride would otherwise pass the check and the reads would leave the buffer)
        if (count > max_accessor_elements)
        {
            throw std::runtime_error("an accessor of " + std::to_string(count) + " elements, more than this loader reads");
        }
        const json* view_index = a.find("bufferView");
        if (!view_index)
        {
            // no buffer view: zeros (glTF 2.0 section 3.6.2)
            return std::vector<double>(count * static_cast<std::size_t>(components), 0.0);
        }
        const json_array& views = array_or_empty(root_, "bufferViews");
        const std::size_t vi = index_of(view_index, "bufferView");
        if (vi >= views.size())
        {
            throw std::runtime_error("a bufferView index past the buffer views");
        }
        const json& v = views[vi];
        const std::size_t buffer = index_of(v.find("buffer"), "buffer");
        if (buffer >= buffers_.size())
        {
            throw std::runtime_error("a buffer index past the buffers");
        }
        const std::size_t buffer_size = buffers_[buffer].size();
        const std::size_t view_offset = v.find("byteOffset") ? index_of(v.find("byteOffset"), "byteOffset") : 0;
        const std::size_t view_length = index_of(v.find("byteLength"), "bufferView byteLength");
        const std::size_t element = size * static_cast<std::size_t>(components);
        const bool strided = v.find("byteStride") != nullptr;
        const std::size_t stride = strided ? index_of(v.find("byteStride"), "byteStride") : element;
        const std::size_t offset = a.find("byteOffset") ? index_of(a.find("byteOffset"), "byteOffset") : 0;
        if (view_offset > buffer_size || view_length > buffer_size - view_offset)
        {
            throw std::runtime_error("a buffer view past its buffer");
        }
        // glTF 2.0 section 5.11: byteStride in [4, 252], and no smaller than an element
        if (strided && (stride < 4 || stride > 252))
        {
            throw std::runtime_error("a buffer view of byteStride " + std::to_string(stride) + ", outside [4, 252]");
        }
        if (stride < element)
        {
            throw std::runtime_error("a buffer view whose byteStride is smaller than its accessor's elements");
        }
        if (count > 0 && (offset > view_length || element > view_length - offset ||
                          count - 1 > (view_length - offset - element) / stride))
        {
            throw std::runtime_error("an accessor past its buffer view");
        }
        std::vector<double> out(count * static_cast<std::size_t>(components), 0.0);
        const uint8_t* base = buffers_[buffer].data() + view_offset + offset;
        for (std::size_t e = 0; e < count; ++e)
        {
            for (int c = 0; c < components; ++c)
            {
                const uint8_t* at = base + e * stride + static_cast<std::size_t>(c) * size;
                double x = 0;
                switch (component_type)
                {
                case 5120: x = static_cast<int8_t>(at[0]); break;
                case 5121: x = at[0]; break;
                case 5122: { int16_t y; std::memcpy(&y, at, 2); x = y; break; }
                case 5123: { uint16_t y; std::memcpy(&y, at, 2); x = y; break; }
                case 5125: { uint32_t y; std::memcpy(&y, at, 4); x = y; break; }
                default: { float y; std::memcpy(&y, at, 4); x = y; break; }
                }
                out[e * static_cast<std::size_t>(components) + static_cast<std::size_t>(c)] = x;
            }
        }
        return out;
    }

    // The mesh of a primitive with triangles, made once and shared by every node that holds it; -1 for none
    int primitive(std::size_t mesh_index, std::size_t p)
    {
        const auto key = std::make_pair(mesh_index, p);
        if (const auto it = primitive_mesh_.find(key); it != primitive_mesh_.end())
        {
            return it->second;
        }
        const json& prim = array_or_empty(root_, "meshes").at(mesh_index).find("primitives")->array().at(p);
        const int mode = prim.find("mode") ? static_cast<int>(index_of(prim.find("mode"), "mode")) : 4;
        int result = -1;
        if (mode < 4)
        {
            model_.warnings.push_back("a primitive of points and lines skipped: only triangles are rendered");
        }
        else if (mode > 6)
        {
            throw std::runtime_error("a primitive of mode " + std::to_string(mode));
        }
        else
        {
            const json* attributes = prim.find("attributes");
            const json* position = attributes ? attributes->find("POSITION") : nullptr;
            if (!position)
            {
                throw std::runtime_error("a primitive without POSITION");
            }
            int components = 0, type = 0;
            const std::vector<double> positions = read_accessor(index_of(position, "POSITION"), components, type);
            if (components != 3 || type != 5126)
            {
                throw std::runtime_error("positions that aren't three floats");
            }
            const std::size_t vertices = positions.size() / 3;
            std::vector<uint32_t> order;
            if (const json* indices = prim.find("indices"))
            {
                int c = 0, t = 0;
                const std::vector<double> raw = read_accessor(index_of(indices, "indices"), c, t);
                if (c != 1 || (t != 5121 && t != 5123 && t != 5125))
                {
                    throw std::runtime_error("indices that aren't unsigned integers");
                }
                for (const double x : raw)
                {
                    if (x >= static_cast<double>(vertices))
                    {
                        throw std::runtime_error("an index past the vertices (" + std::to_string(static_cast<uint64_t>(x)) +
                                                 " of " + std::to_string(vertices) + ")");
                    }
                    order.push_back(static_cast<uint32_t>(x));
                }
            }
            else
            {
                for (std::size_t k = 0; k < vertices; ++k)
                {
                    order.push_back(static_cast<uint32_t>(k));
                }
            }
            gltf_mesh m;
            m.positions.assign(positions.begin(), positions.end());
            // triangles (4), strips (5: every other one turned to keep the winding) and fans (6)
            const std::size_t n = order.size();
            for (std::size_t k = 0; mode == 4 && k + 3 <= n; k += 3)
            {
                m.indices.insert(m.indices.end(), {order[k], order[k + 1], order[k + 2]});
            }
            for (std::size_t k = 0; mode == 5 && k + 2 < n; ++k)
            {
                if (k % 2 == 0)
                {
                    m.indices.insert(m.indices.end(), {order[k], order[k + 1], order[k + 2]});
                }
                else
                {
                    m.indices.insert(m.indices.end(), {order[k + 1], order[k], order[k + 2]});
                }
            }
            for (std::size_t k = 1; mode == 6 && k + 1 < n; ++k)
            {
                m.indices.insert(m.indices.end(), {order[0], order[k], order[k + 1]});
            }
            if (mode == 4 && n % 3 != 0)
            {
                model_.warnings.push_back("a triangle list whose length isn't a multiple of 3: the last indices skipped");
            }
            if (const json* material = prim.find("material"))
            {
                m.material = static_cast<int>(index_of(material, "material"));
                if (static_cast<std::size_t>(m.material) >= model_.materials.size())
                {
                    throw std::runtime_error("a material index past the materials");
                }
            }
            if (!m.indices.empty())
            {
                result = static_cast<int>(model_.meshes.size());
                model_.meshes.push_back(std::move(m));
            }
        }
        primitive_mesh_[key] = result;
        return result;
    }

    void place_node(std::size_t index, const mat4& parent, std::vector<bool>& on_path)
    {
        const json_array& nodes = array_or_empty(root_, "nodes");
        if (index >= nodes.size())
        {
            throw std::runtime_error("a node index past the nodes");
        }
        if (on_path[index])
        {
            throw std::runtime_error("the node hierarchy has a cycle (node " + std::to_string(index) + ")");
        }
        on_path[index] = true;
        const json& node = nodes[index];
        const mat4 world = multiply(parent, local_transform(node));
        if (const json* mesh = node.find("mesh"))
        {
            const std::size_t mi = index_of(mesh, "mesh");
            const json_array& meshes = array_or_empty(root_, "meshes");
            if (mi >= meshes.size())
            {
                throw std::runtime_error("a mesh index past the meshes");
            }
            const json* prims = meshes[mi].find("primitives");
            for (std::size_t p = 0; prims && p < prims->array().size(); ++p)
            {
                const int m = primitive(mi, p);
                if (m < 0)
                {
                    continue;
                }
                gltf_placement placement;
                placement.mesh = m;
                for (int row = 0; row < 3; ++row)
                {
                    for (int col = 0; col < 4; ++col)
                    {
                        placement.object_to_world[row * 4 + col] = world[col * 4 + row];
                    }
                }
                model_.placements.push_back(placement);
            }
        }
        if (node.find("skin") || node.find("weights"))
        {
            model_.warnings.push_back("skins and morph targets are not applied: the meshes render in their rest pose");
        }
        if (const json* children = node.find("children"))
        {
            for (const json& c : children->array())
            {
                place_node(index_of(&c, "child node"), world, on_path);
            }
        }
        on_path[index] = false;
    }

    // The scene's root nodes (the default scene, else the first, else every node no other node lists as a child)
    void place_scene()
    {
        const json_array& nodes = array_or_empty(root_, "nodes");
        std::vector<std::size_t> roots;
        const json_array& scenes = array_or_empty(root_, "scenes");
        if (!scenes.empty())
        {
            const std::size_t s = root_.find("scene") ? index_of(root_.find("scene"), "scene") : 0;
            if (s >= scenes.size())
            {
                throw std::runtime_error("a scene index past the scenes");
            }
            if (const json* list = scenes[s].find("nodes"))
            {
                for (const json& n : list->array())
                {
                    roots.push_back(index_of(&n, "scene node"));
                }
            }
        }
        else
        {
            std::vector<bool> child(nodes.size(), false);
            for (const json& n : nodes)
            {
                if (const json* c = n.find("children"))
                {
                    for (const json& k : c->array())
                    {
                        const std::size_t i = index_of(&k, "child node");
                        if (i < child.size())
                        {
                            child[i] = true;
                        }
                    }
                }
            }
            for (std::size_t i = 0; i < nodes.size(); ++i)
            {
                if (!child[i])
                {
                    roots.push_back(i);
                }
            }
        }
        std::vector<bool> on_path(nodes.size(), false);
        for (const std::size_t r : roots)
        {
            place_node(r, identity(), on_path);
        }
        // warnings once each
        std::vector<std::string> unique;
        for (const std::string& w : model_.warnings)
        {
            if (std::find(unique.begin(), unique.end(), w) == unique.end())
            {
                unique.push_back(w);
            }
        }
        model_.warnings = std::move(unique);
    }
};
}

gltf_model load_gltf(const std::string& path)
{
    try
    {
        return loader(path).load();
    }
    catch (const std::runtime_error& e)
    {
        throw std::runtime_error(path + ": " + e.what());
    }
    catch (const std::out_of_range&)
    {
        throw std::runtime_error(path + ": an array of the glTF JSON is shorter than an index into it");
    }
}

bool is_model_path(const std::string& name)
{
    auto ends = [&](const char* suffix) {
        const std::size_t n = std::strlen(suffix);
        if (name.size() < n)
        {
            return false;
        }
        for (std::size_t k = 0; k < n; ++k)
        {
            if (std::tolower(static_cast<unsigned char>(name[name.size() - n + k])) != suffix[k])
            {
                return false;
            }
        }
        return true;
    };
    return ends(".gltf") || ends(".glb");
}
}
