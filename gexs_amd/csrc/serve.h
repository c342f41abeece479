// serve.h -- continuous-batching request server over a resident line list.
#pragma once

namespace gxs {

class Server {
public:
    virtual ~Server() = default;
};

}  // namespace gxs
