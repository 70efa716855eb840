"""gRPC front end of the encoding service: ``encoder.Encoder/EncodeBlob``
(grpc/proto/encoder.proto:5-21, grpc/src/service.rs:47-64, grpc/src/lib.rs:11-33).

The method is served through grpcio's generic handlers with the protobuf codec of
``wire`` (no generated stubs), with the reference's 1 GiB message limit in both
directions.  A request that ``EncoderService`` rejects with TooLargeBlob ends with
status UNKNOWN and the error's Debug text as its message, as tonic's
``Status::new(Code::Unknown, e)`` reports it (service.rs:59-61,
crates/encoder/src/encoder/error.rs:23-25).  A served request whose client goes away
before it reaches the GPU is withdrawn from the server's queue.  grpcio is imported when a server or
client is made, not with the package.
"""
from concurrent import futures

from . import wire

SERVICE = "encoder.Encoder"
METHOD = "EncodeBlob"
PATH = "/%s/%s" % (SERVICE, METHOD)


def _options(limit):
    return [("grpc.max_receive_message_length", limit), ("grpc.max_send_message_length", limit)]


def serve(service, address="127.0.0.1:0", max_workers=4, message_limit=wire.MESSAGE_SIZE_LIMIT):
    """Starts a server for ``service`` (an EncoderService, or anything with its
    ``encode_blob(data, require_data, alive) -> reply bytes``) on ``address``.  Returns
    (server, bound port); stop it with ``server.stop(grace)``."""
    import grpc

    from . import RequestCancelled, TooLargeBlob

    def encode_blob(request, context):
        data, require_data = wire.decode_request(request)
        try:
            return service.encode_blob(data, require_data, alive=context.is_active)
        except TooLargeBlob as e:
            context.abort(grpc.StatusCode.UNKNOWN, str(e))
        except RequestCancelled as e:  # the client is gone; the status is never read
            context.abort(grpc.StatusCode.CANCELLED, str(e))

    handler = grpc.unary_unary_rpc_method_handler(encode_blob, request_deserializer=None, response_serializer=None)
    server = grpc.server(futures.ThreadPoolExecutor(max_workers=max_workers), options=_options(message_limit))
    server.add_generic_rpc_handlers((grpc.method_handlers_generic_handler(SERVICE, {METHOD: handler}),))
    port = server.add_insecure_port(address)
    if port == 0:
        raise RuntimeError("cannot bind %s" % address)
    server.start()
    return server, port


class EncoderClient:
    """Client of ``encoder.Encoder/EncodeBlob`` (the call a 0G node makes)."""

    def __init__(self, target, message_limit=wire.MESSAGE_SIZE_LIMIT):
        import grpc

        self._channel = grpc.insecure_channel(target, options=_options(message_limit))
        self._call = self._channel.unary_unary(PATH, request_serializer=None, response_deserializer=None)

    def encode_blob_raw(self, data, require_data=False, timeout=None):
        """-> the reply's protobuf bytes."""
        return self._call(wire.encode_request(data, require_data), timeout=timeout)

    def encode_blob(self, data, require_data=False, timeout=None):
        """-> the decoded EncodeBlobReply fields (wire.decode_reply)."""
        return wire.decode_reply(self.encode_blob_raw(data, require_data, timeout))

    def close(self):
        self._channel.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
