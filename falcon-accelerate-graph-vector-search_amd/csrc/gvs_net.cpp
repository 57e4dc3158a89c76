// TCP front end with the wire format of the reference's networked search
// (networked_FPGA/CPU_programs/README.md:40-72).  The reference's client
// (networked_FPGA/CPU_programs/CPU_client.cpp:450-620) keeps two
// connections to the search node: it sends batches on one (it connects to
// the node's query port) and receives results on the other (it listens, the
// node connects).  The node side here stands where the FPGA's network kernel
// does (networked_FPGA/kernel/user_krnl/FPGA_inter_query_v1_3/src/hls/
// FPGA_inter_query_v1_3.cpp:64-100, :187-230; the software stand-in is
// networked_FPGA/CPU_programs/FPGA_simulator.cpp):
//
//   client -> node, per batch: a 64-byte header whose first int32 is the
//     batch size (<= 0 ends the session), then batch-size queries of
//     ceil(D/16) 64-byte packets each (fp32, zero padded);
//   node -> client, per query in arrival order: a 64-byte header whose
//     first int32 is ef, then ceil(ef/16) packets of int32 ids and
//     ceil(ef/16) packets of fp32 distances, ascending (zero padded).
//
// Queries go to the GPU serving loop (gvs_server_*) as they arrive, so the
// searches of one batch, and of the next batches, overlap on the device.  A
// request whose visited table overflowed is searched again on the batch
// path (gvs_search_host retries with larger tables), so every reply is the
// exact search result.  Host code only: it reaches the device through the
// C ABI of the serving loop.
#include <arpa/inet.h>
#include <netdb.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <sys/socket.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gvs.h"
#include "host_graph.hpp"

namespace gvs {
int server_search_setup(gvs_server *s, gvs_index **ix, gvs_params *p, int *max_k, int *max_ef);
}

namespace {

constexpr int64_t PACKET = 64;      // bytes per 512-bit network packet (BYTES_PER_AXI)
constexpr int64_t PER_PACKET = 16;  // int32 / fp32 values per packet

int64_t packets(int64_t values) { return (values + PER_PACKET - 1) / PER_PACKET; }

using Clock = std::chrono::steady_clock;

struct Fd {
    int fd = -1;
    ~Fd()
    {
        if (fd >= 0) ::close(fd);
    }
};

bool read_full(int fd, void *buf, size_t n)
{
    char *p = (char *)buf;
    while (n > 0) {
        const ssize_t r = ::recv(fd, p, n, 0);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
    }
    return true;
}

bool write_full(int fd, const void *buf, size_t n)
{
    const char *p = (const char *)buf;
    while (n > 0) {
        const ssize_t r = ::send(fd, p, n, MSG_NOSIGNAL);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
    }
    return true;
}

void no_delay(int fd)
{
    int one = 1;
    (void)setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof one);
}

// Connect to host:port, retrying until the deadline (the client may start
// listening after the node starts).
int connect_retry(const char *host, int port, Clock::time_point deadline)
{
    addrinfo hints;
    std::memset(&hints, 0, sizeof hints);
    hints.ai_family = AF_INET;
    hints.ai_socktype = SOCK_STREAM;
    const std::string ps = std::to_string(port);
    for (;;) {
        addrinfo *res = nullptr;
        if (getaddrinfo(host, ps.c_str(), &hints, &res) == 0 && res) {
            const int fd = ::socket(res->ai_family, res->ai_socktype, res->ai_protocol);
            if (fd >= 0 && ::connect(fd, res->ai_addr, res->ai_addrlen) == 0) {
                freeaddrinfo(res);
                no_delay(fd);
                return fd;
            }
            if (fd >= 0) ::close(fd);
            freeaddrinfo(res);
        }
        if (Clock::now() >= deadline) return -1;
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
    }
}

// One query between its arrival and its reply.
struct Pending {
    std::vector<float> query;  // kept for a batch-path re-run
    bool done = false;
    bool retry = false;        // the serving loop reported a visited-table overflow
    std::vector<int64_t> labels;
    std::vector<float> dists;
};

}  // namespace

extern "C" int gvs_net_frame_bytes(int32_t dim, int32_t ef, int64_t out[3])
{
    if (dim < 1 || ef < 1 || !out) {
        gvs::set_error("gvs_net_frame_bytes: invalid argument");
        return GVS_ERR_ARG;
    }
    out[0] = PACKET;                             // batch header
    out[1] = packets(dim) * PACKET;              // one query
    out[2] = PACKET + 2 * packets(ef) * PACKET;  // one reply: header, ids, distances
    return GVS_OK;
}

extern "C" int gvs_net_serve(gvs_server *s, const gvs_net_config *cfg, int32_t *bound_port, int64_t stats[6])
{
    if (!s || !cfg || cfg->ef < 1 || cfg->dim < 1 || cfg->c2f_port < 0 || cfg->c2f_port > 65535 ||
        cfg->f2c_port < 1 || cfg->f2c_port > 65535 || !cfg->f2c_host) {
        gvs::set_error("gvs_net_serve: invalid configuration");
        return GVS_ERR_ARG;
    }
    gvs_index *ix = nullptr;
    gvs_params rp;
    int max_k = 0, max_ef = 0;
    if (gvs::server_search_setup(s, &ix, &rp, &max_k, &max_ef) != GVS_OK) {
        gvs::set_error("gvs_net_serve: invalid server");
        return GVS_ERR_ARG;
    }
    const int ef = cfg->ef;
    if (ef > max_k || ef > max_ef) {
        gvs::set_error("gvs_net_serve: ef %d exceeds the server's max_k %d / max_ef %d (replies hold ef results)", ef,
                       max_k, max_ef);
        return GVS_ERR_ARG;
    }
    const int64_t dim = cfg->dim;
    const int64_t q_bytes = packets(dim) * PACKET;
    const int64_t ldq = q_bytes / 4;  // floats per query on the wire; the first dim are the vector
    const int64_t id_bytes = packets(ef) * PACKET;
    const int64_t reply_bytes = PACKET + 2 * id_bytes;
    const int timeout_ms = cfg->timeout_ms > 0 ? cfg->timeout_ms : 10000;
    const Clock::time_point t_open = Clock::now();
    rp.k = ef;
    rp.ef = ef;

    // query port: listen, report the bound port, accept the client
    Fd lfd;
    lfd.fd = ::socket(AF_INET, SOCK_STREAM, 0);
    if (lfd.fd < 0) {
        gvs::set_error("gvs_net_serve: socket: %s", std::strerror(errno));
        return GVS_ERR_IO;
    }
    int one = 1;
    (void)setsockopt(lfd.fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
    sockaddr_in addr;
    std::memset(&addr, 0, sizeof addr);
    addr.sin_family = AF_INET;
    addr.sin_addr.s_addr = htonl(INADDR_ANY);
    addr.sin_port = htons((uint16_t)cfg->c2f_port);
    if (::bind(lfd.fd, (sockaddr *)&addr, sizeof addr) != 0 || ::listen(lfd.fd, 1) != 0) {
        gvs::set_error("gvs_net_serve: bind/listen on port %d: %s", cfg->c2f_port, std::strerror(errno));
        return GVS_ERR_IO;
    }
    socklen_t alen = sizeof addr;
    (void)getsockname(lfd.fd, (sockaddr *)&addr, &alen);
    if (bound_port) __atomic_store_n(bound_port, (int32_t)ntohs(addr.sin_port), __ATOMIC_RELEASE);

    // result port: connect to the client's listener
    Fd ofd;
    ofd.fd = connect_retry(cfg->f2c_host, cfg->f2c_port, t_open + std::chrono::milliseconds(timeout_ms));
    if (ofd.fd < 0) {
        gvs::set_error("gvs_net_serve: cannot connect to %s:%d", cfg->f2c_host, cfg->f2c_port);
        return GVS_ERR_IO;
    }
    pollfd pfd{lfd.fd, POLLIN, 0};
    const int left = (int)std::max<int64_t>(
        1, timeout_ms - std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - t_open).count());
    if (::poll(&pfd, 1, left) != 1) {
        gvs::set_error("gvs_net_serve: no client connected to port %d", (int)ntohs(addr.sin_port));
        return GVS_ERR_IO;
    }
    Fd ifd;
    ifd.fd = ::accept(lfd.fd, nullptr, nullptr);
    if (ifd.fd < 0) {
        gvs::set_error("gvs_net_serve: accept: %s", std::strerror(errno));
        return GVS_ERR_IO;
    }
    no_delay(ifd.fd);

    // Requests in arrival order: the receiver appends (submitting under the
    // lock, so a completion always finds its entry), the sender completes
    // entries and replies from the front.
    std::mutex mu;
    std::deque<Pending> order;
    std::unordered_map<int64_t, int64_t> where;  // request id -> arrival number
    int64_t first = 0;                           // arrival number of order.front()
    bool recv_done = false, failed = false;
    std::string fail_msg;
    int64_t n_batches = 0, n_queries = 0, n_retried = 0, n_sent = 0, bytes_in = 0, bytes_out = 0;
    // the first failure ends the session; the query socket is shut down so
    // that a receiver blocked on it returns
    const auto fail = [&](const std::string &m) {
        std::lock_guard<std::mutex> g(mu);
        if (!failed) fail_msg = m;
        failed = true;
        (void)::shutdown(ifd.fd, SHUT_RD);
    };

    std::thread sender([&]() {
        const int B = 256;
        std::vector<int64_t> rid(B), lab((size_t)B * max_k);
        std::vector<float> dis((size_t)B * max_k);
        std::vector<int32_t> st(B);
        std::vector<char> reply((size_t)reply_bytes);
        std::vector<int64_t> lab1(ef);
        std::vector<float> dis1(ef);
        for (;;) {
            const int got = gvs_server_poll_status(s, B, rid.data(), lab.data(), dis.data(), nullptr, nullptr,
                                                   st.data());
            if (got < 0) {
                fail(std::string("poll: ") + gvs_last_error());
                return;
            }
            std::vector<int64_t> redo;
            {
                std::lock_guard<std::mutex> g(mu);
                for (int j = 0; j < got; j++) {
                    auto it = where.find(rid[j]);
                    if (it == where.end()) continue;  // not one of this session's requests
                    Pending &p = order[(size_t)(it->second - first)];
                    if (st[j] == GVS_ERR_OVERFLOW) {
                        p.retry = true;
                        redo.push_back(it->second);
                    } else {
                        p.labels.assign(lab.begin() + (size_t)j * max_k, lab.begin() + (size_t)j * max_k + ef);
                        p.dists.assign(dis.begin() + (size_t)j * max_k, dis.begin() + (size_t)j * max_k + ef);
                        p.done = true;
                    }
                    where.erase(it);
                }
            }
            // re-run overflowed requests on the batch path (exact, with
            // retries on larger tables); entries stay in place until replied
            for (const int64_t a : redo) {
                std::vector<float> q;
                {
                    std::lock_guard<std::mutex> g(mu);
                    q = order[(size_t)(a - first)].query;
                }
                const int rc = gvs_search_host(ix, q.data(), 1, dim, nullptr, &rp, lab1.data(), dis1.data(), nullptr);
                if (rc != GVS_OK) {
                    fail(std::string("batch-path re-run: ") + gvs_last_error());
                    return;
                }
                std::lock_guard<std::mutex> g(mu);
                Pending &p = order[(size_t)(a - first)];
                p.labels = lab1;
                p.dists = dis1;
                p.done = true;
                n_retried++;
            }
            // reply to every finished request at the front, in order
            int sent = 0;
            for (;;) {
                Pending p;
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (failed) return;
                    if (order.empty() || !order.front().done) break;
                    p = std::move(order.front());
                    order.pop_front();
                    first++;
                }
                std::memset(reply.data(), 0, reply.size());
                const int32_t hdr = ef;
                std::memcpy(reply.data(), &hdr, 4);
                int32_t *ids = (int32_t *)(reply.data() + PACKET);
                float *ds = (float *)(reply.data() + PACKET + id_bytes);
                for (int i = 0; i < ef; i++) {
                    const int64_t l = p.labels[i];
                    if (l > INT32_MAX || l < -1) {
                        fail("label does not fit the reply's int32 id field");
                        return;
                    }
                    ids[i] = (int32_t)l;
                    ds[i] = p.dists[i];
                }
                if (!write_full(ofd.fd, reply.data(), reply.size())) {
                    fail(std::string("result connection: ") + std::strerror(errno));
                    return;
                }
                std::lock_guard<std::mutex> g(mu);
                n_sent++;
                bytes_out += reply_bytes;
                sent++;
            }
            if (got == 0 && sent == 0) {
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (failed || (recv_done && order.empty())) return;
                }
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    });

    // receiver: batches from the query connection into the serving loop
    std::vector<char> hdr((size_t)PACKET);
    std::vector<float> qbuf;
    std::vector<int32_t> ks, efs;
    std::vector<int64_t> ids;
    for (;;) {
        if (!read_full(ifd.fd, hdr.data(), hdr.size())) {
            fail("query connection closed before the end-of-session header");
            break;
        }
        int32_t bs = 0;
        std::memcpy(&bs, hdr.data(), 4);
        {
            std::lock_guard<std::mutex> g(mu);
            bytes_in += PACKET;
            if (failed) break;
        }
        if (bs <= 0) break;  // end of session (the reference sends -1)
        qbuf.resize((size_t)bs * ldq);
        if (!read_full(ifd.fd, qbuf.data(), (size_t)bs * q_bytes)) {
            fail("query connection closed inside a batch");
            break;
        }
        ks.assign(bs, ef);
        efs.assign(bs, ef);
        ids.resize(bs);
        int64_t i = 0;
        bool stop = false;
        while (i < bs && !stop) {
            int64_t m;
            {
                std::lock_guard<std::mutex> g(mu);
                if (failed) {
                    stop = true;
                    break;
                }
                m = gvs_server_submit_batch(s, qbuf.data() + i * ldq, bs - i, ldq, ks.data() + i, efs.data() + i,
                                            ids.data() + i);
                if (m < 0) {
                    fail_msg = std::string("submit: ") + gvs_last_error();
                    failed = true;
                    stop = true;
                    break;
                }
                // m == 0: every device slot is busy until the sender's
                // polls collect completions
                for (int64_t j = i; j < i + m; j++) {
                    Pending p;
                    p.query.assign(qbuf.begin() + j * ldq, qbuf.begin() + j * ldq + dim);
                    where[ids[j]] = first + (int64_t)order.size();
                    order.push_back(std::move(p));
                }
                bytes_in += m * q_bytes;
                n_queries += m;
            }
            i += m;
            if (i < bs) std::this_thread::sleep_for(std::chrono::microseconds(10));  // slots free as replies go out
        }
        if (stop) break;
        std::lock_guard<std::mutex> g(mu);
        n_batches++;
    }
    {
        std::lock_guard<std::mutex> g(mu);
        recv_done = true;
    }
    sender.join();
    if (stats) {
        stats[0] = n_batches;
        stats[1] = n_queries;
        stats[2] = n_sent;
        stats[3] = n_retried;
        stats[4] = bytes_in;
        stats[5] = bytes_out;
    }
    if (failed) {
        gvs::set_error("gvs_net_serve: %s", fail_msg.c_str());
        return GVS_ERR_IO;
    }
    return GVS_OK;
}
